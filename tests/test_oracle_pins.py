"""Pins of the oracle's high-order stages by what the mathematics fixes (CPU only).

* ``fill_up_dn_grad`` (fesom_jax/tracer_adv.py:344-368): with element gradients set to
  the element's own id, every slot inside the levels where all triangles around both
  edge ends are wet must read the up/down triangle's id, and every other slot the node
  average at that end.
* ``adv_flux_hor_ho`` (MFCT, fesom_jax/tracer_adv.py:196-225): for a tracer linear in
  the metric coordinates of a flat mesh, the element gradients are exact, so both edge
  means equal the edge-midpoint value and the HO flux is ``-vflux * (T1 + T2) / 2``.
* ``adv_flux_ver_ho`` (QR4C, fesom_jax/tracer_adv.py:241-295): for a column linear in
  depth, both interior interface means equal T at the interface depth ``zbar[k]``, the
  centred interfaces take the mean of the two layers, the surface takes its own layer,
  the bottom 0 — with the static stencil and with ``Z3d / zbar3``.

Each planted mistake (wrong MFCT denominator, dropped weight of the centred QR4C
gradient, swapped up/down slots) is put into the oracle by monkeypatching and must
make its check fail.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from fesom_jax_amd.mesh import mesh_from_arrays
from fesom_jax_amd.meshgen import derive_bundle, default_zbar
from oracle import fct_numpy as O

R = 6367500.0


# --------------------------------------------------------------------------
# a flat mesh: jittered grid of nodes in (lon, lat) radians, elem_cos = 1
# --------------------------------------------------------------------------
def _flat_mesh(nx=11, ny=9, nl=12, h=0.01, seed=5):
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    lon = 0.3 + h * ii.ravel() + rng.uniform(-0.2, 0.2, ii.size) * h
    lat = 0.1 + h * jj.ravel() + rng.uniform(-0.2, 0.2, ii.size) * h
    node = lambda i, j: i * ny + j
    tris = []
    for i in range(nx - 1):
        for j in range(ny - 1):
            a, b, c, d = node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)
            if (i + j) % 2:
                tris += [(a, b, c), (a, c, d)]
            else:
                tris += [(a, b, d), (b, c, d)]
    en = np.array(tris, np.int64)
    # clockwise in the (lon, lat) plane
    cr = ((lon[en[:, 1]] - lon[en[:, 0]]) * (lat[en[:, 2]] - lat[en[:, 0]])
          - (lat[en[:, 1]] - lat[en[:, 0]]) * (lon[en[:, 2]] - lon[en[:, 0]]))
    ccw = cr > 0
    en[ccw, 1], en[ccw, 2] = en[ccw, 2].copy(), en[ccw, 1].copy()
    # edges: the triangle whose clockwise half-edge runs n2 -> n1 lies left of n1 -> n2
    half = {}
    for t, (p, q, r) in enumerate(en):
        for u, v in ((p, q), (q, r), (r, p)):
            half[(int(u), int(v))] = t
    edges, etri, seen = [], [], set()
    for (u, v), t in half.items():
        key = (min(u, v), max(u, v))
        if key in seen:
            continue
        seen.add(key)
        if (v, u) in half:                      # interior: n1 < n2
            n1, n2 = key
            left, right = half[(n2, n1)], half[(n1, n2)]
            edges.append((n1, n2)); etri.append((left, right))
        else:                                   # boundary: the one triangle on the left
            edges.append((v, u)); etri.append((t, -1))
    edges = np.array(edges, np.int32); etri = np.array(etri, np.int32)
    order = np.argsort(etri[:, 1] < 0, kind="stable")     # interior edges first
    edges, etri = edges[order], etri[order]
    N = lon.size
    zbar = default_zbar(nl)
    node_lev = rng.integers(nl // 2, nl + 1, N).astype(np.int32)
    nlevels = node_lev[en].min(axis=1).astype(np.int32)
    nln = np.zeros(N, np.int32)
    np.maximum.at(nln, en.reshape(-1), np.repeat(nlevels, 3))
    coord = np.stack([lon, lat], axis=1)
    coast = np.zeros(N, np.int32)
    coast[edges[etri[:, 1] < 0].reshape(-1)] = 1
    arrays, meta = derive_bundle(coord, coord.copy(), coast, en.astype(np.int32), edges,
                                 etri, nlevels, nln, zbar, -zbar[node_lev - 1])
    # flat metric: x = R*lon, y = R*lat; gradient coefficients from a 3x3 solve of the
    # linear interpolant on each triangle (independent of the closed form)
    arrays["elem_cos"] = np.ones(en.shape[0])
    X, Y = R * lon, R * lat
    gs = np.zeros((en.shape[0], 6))
    for t, tri in enumerate(arrays["elem_nodes"]):
        A = np.stack([np.ones(3), X[tri], Y[tri]], axis=1)
        inv = np.linalg.inv(A)                  # rows 1, 2: d/dx, d/dy of the basis
        gs[t, 0:3] = inv[1]
        gs[t, 3:6] = inv[2]
    arrays["gradient_sca"] = gs
    return mesh_from_arrays(arrays, meta), X, Y


@pytest.fixture(scope="module")
def flat():
    return _flat_mesh()


def _linear_tracer(mesh, X, Y, seed=1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(5.0, 20.0, mesh.nl)
    gx = rng.uniform(-2e-5, 2e-5, mesh.nl)
    gy = rng.uniform(-2e-5, 2e-5, mesh.nl)
    return c[None, :] + gx[None, :] * X[:, None] + gy[None, :] * Y[:, None]


def _flow(mesh, seed=2):
    rng = np.random.default_rng(seed)
    uv = rng.uniform(-0.3, 0.3, (mesh.elem2D, mesh.nl, 2))
    helem = np.where(mesh.elem_layer_mask, rng.uniform(5.0, 50.0, (mesh.elem2D, mesh.nl)), 0.0)
    return uv, helem


# --------------------------------------------------------------------------
# MFCT (a7 -> a8 -> a9 -> a10)
# --------------------------------------------------------------------------
def _mfct_residual(mesh, X, Y):
    T = _linear_tracer(mesh, X, Y)
    uv, helem = _flow(mesh)
    tr_xy = O.tracer_gradient_elements(mesh, T)
    eud = O.fill_up_dn_grad(mesh, tr_xy)
    ho = O.adv_flux_hor_ho(mesh, uv, helem, T, eud)
    vflux = O.edge_vflux(mesh, uv, helem)
    n1, n2 = mesh.edges[:, 0], mesh.edges[:, 1]
    mid = 0.5 * (T[n1] + T[n2])
    want = -vflux * mid
    scale = np.max(np.abs(want))
    return float(np.max(np.abs(ho - want)) / scale), int(np.sum(vflux != 0.0))


def test_mfct_linear_field_gives_midpoint_flux(flat):
    mesh, X, Y = flat
    err, nflux = _mfct_residual(mesh, X, Y)
    assert nflux > 0.5 * mesh.edge2D * (mesh.nl // 2)
    assert err < 1e-12, err


def test_planted_mfct_denominator_breaks_midpoint(flat, monkeypatch):
    mesh, X, Y = flat
    monkeypatch.setattr(O, "_MFCT_DEN", 4.0)
    err, _ = _mfct_residual(mesh, X, Y)
    assert err > 1e-4, err


def test_planted_mfct_difference_weight_breaks_midpoint(flat, monkeypatch):
    mesh, X, Y = flat
    monkeypatch.setattr(O, "_MFCT_DIFF_W", 1.0)
    err, _ = _mfct_residual(mesh, X, Y)
    assert err > 1e-4, err


def test_mfct_gradient_terms_enter_with_their_metric(flat):
    """The same check with the gradient slots zeroed fails: the edge means are carried
    by the metric gradient terms, not by the difference term alone."""
    mesh, X, Y = flat
    T = _linear_tracer(mesh, X, Y)
    uv, helem = _flow(mesh)
    eud = np.zeros((mesh.edge2D, mesh.nl, 4))
    ho = O.adv_flux_hor_ho(mesh, uv, helem, T, eud)
    vflux = O.edge_vflux(mesh, uv, helem)
    mid = 0.5 * (T[mesh.edges[:, 0]] + T[mesh.edges[:, 1]])
    assert np.max(np.abs(ho + vflux * mid)) / np.max(np.abs(vflux * mid)) > 1e-4


# --------------------------------------------------------------------------
# fill_up_dn_grad (a9): slot selection and level range
# --------------------------------------------------------------------------
def _all_fan_wet(mesh):
    """[D, nl]: every triangle around either end of the edge is wet at level k (built
    from the node -> element lists and the element layer mask)."""
    off, nie = mesh.nod_in_elem2D_offsets, mesh.nod_in_elem2D
    node_all = np.ones((mesh.nod2D, mesh.nl), bool)
    for n in range(mesh.nod2D):
        node_all[n] = mesh.elem_layer_mask[nie[off[n]:off[n + 1]]].all(axis=0)
    return node_all[mesh.edges[:, 0]] & node_all[mesh.edges[:, 1]]


def _check_slots(mesh):
    E, nl = mesh.elem2D, mesh.nl
    tr_xy = np.zeros((E, nl, 2))
    ids = np.arange(E, dtype=np.float64)[:, None]
    tr_xy[..., 0] = np.where(mesh.elem_layer_mask, ids + 0.25, 0.0)
    tr_xy[..., 1] = np.where(mesh.elem_layer_mask, -ids - 0.5, 0.0)
    eud = O.fill_up_dn_grad(mesh, tr_xy)
    gnode = O.node_avg_grad(mesh, tr_xy)
    up, dn = mesh.edge_up_dn_tri[:, 0], mesh.edge_up_dn_tri[:, 1]
    use = _all_fan_wet(mesh) & ((up >= 0) & (dn >= 0))[:, None]
    n1, n2 = mesh.edges[:, 0], mesh.edges[:, 1]
    want = np.empty_like(eud)
    bu = np.broadcast_to((up + 0.25)[:, None], use.shape)
    bd = np.broadcast_to((dn + 0.25)[:, None], use.shape)
    want[..., 0] = np.where(use, bu, gnode[n1][..., 0])
    want[..., 1] = np.where(use, bd, gnode[n2][..., 0])
    want[..., 2] = np.where(use, -bu - 0.25, gnode[n1][..., 1])
    want[..., 3] = np.where(use, -bd - 0.25, gnode[n2][..., 1])
    return eud, want, use


@pytest.mark.parametrize("which", ["pi", "flat"])
def test_fill_up_dn_grad_slots(which, pi_mesh, flat):
    mesh = pi_mesh if which == "pi" else flat[0]
    eud, want, use = _check_slots(mesh)
    assert use.any() and (~use).any()
    np.testing.assert_array_equal(eud, want)


def test_planted_swapped_up_dn_slots_fail(pi_mesh, monkeypatch):
    orig = O.fill_up_dn_grad

    def swapped(mesh, tr_xy):
        e = orig(mesh, tr_xy)
        return e[..., [1, 0, 3, 2]]

    monkeypatch.setattr(O, "fill_up_dn_grad", swapped)
    eud, want, use = _check_slots(pi_mesh)
    assert np.any(eud != want)


# --------------------------------------------------------------------------
# QR4C (a11)
# --------------------------------------------------------------------------
def _qr4c_check(mesh, zstar):
    rng = np.random.default_rng(7)
    N, nl = mesh.nod2D, mesh.nl
    alpha = rng.uniform(2.0, 20.0, N)[:, None]
    beta = rng.uniform(-0.01, 0.01, N)[:, None]
    if zstar:
        s = rng.uniform(0.9, 1.1, N)[:, None]
        zbar3 = mesh.zbar[None, :] * s
        Z3d = np.concatenate([0.5 * (zbar3[:, :-1] + zbar3[:, 1:]), zbar3[:, -1:]], axis=1)
        Zc, zb = Z3d, zbar3
    else:
        Z3d = zbar3 = None
        Zc = np.concatenate([mesh.Z, mesh.Z[-1:]])[None, :].repeat(N, 0)
        zb = np.broadcast_to(mesh.zbar[None, :], (N, nl))
    T = alpha + beta * Zc                        # linear in the layer depth
    W = rng.uniform(-1e-4, 1e-4, (N, nl))
    out = O.adv_flux_ver_ho(mesh, W, T, Z3d=Z3d, zbar3=zbar3)
    k = np.arange(nl)[None, :]
    nzmin = (mesh.ulevels_nod2D.astype(np.int64) - 1)[:, None]
    nzmax = (mesh.nlevels_nod2D.astype(np.int64) - 1)[:, None]
    area = mesh.area
    interior = (k >= nzmin + 2) & (k <= nzmax - 2)
    centred = ((k == nzmin + 1) | (k == nzmax - 1)) & ~interior & (k < nzmax)
    Zm1 = np.concatenate([Zc[:, :1], Zc[:, :-1]], axis=1)
    want_int = -(alpha + beta * zb) * W * area
    want_cen = -(alpha + beta * (0.5 * (Zm1 + Zc))) * W * area
    want_srf = -T * W * area
    scale = np.max(np.abs(want_int[interior]))
    r_int = np.max(np.abs(out - want_int)[interior]) / scale
    r_cen = np.max(np.abs(out - want_cen)[centred]) / scale
    srf = k == nzmin
    r_srf = np.max(np.abs(out - want_srf)[srf]) / scale
    below = k >= nzmax
    return r_int, r_cen, r_srf, float(np.max(np.abs(out[below]))), int(interior.sum())


@pytest.mark.parametrize("zstar", [False, True])
def test_qr4c_linear_column(pi_mesh, zstar):
    r_int, r_cen, r_srf, below, nint = _qr4c_check(pi_mesh, zstar)
    assert nint > 1000
    assert r_int < 1e-12, r_int
    assert r_cen < 1e-12, r_cen
    assert r_srf == 0.0
    assert below == 0.0


@pytest.mark.parametrize("zstar", [False, True])
def test_planted_qr4c_centred_weight_fails(pi_mesh, zstar, monkeypatch):
    monkeypatch.setattr(O, "_QR4C_QC_W", 1.0)
    r_int, *_ = _qr4c_check(pi_mesh, zstar)
    assert r_int > 1e-6, r_int


def test_planted_qr4c_denominator_fails(pi_mesh, monkeypatch):
    monkeypatch.setattr(O, "_QR4C_DEN", 2.0)
    r_int, *_ = _qr4c_check(pi_mesh, False)
    assert r_int > 1e-6, r_int


def test_planted_qr4c_stencil_shift_fails(pi_mesh, monkeypatch):
    """A k-2 stencil that repeats k-1 (qd taken over the wrong pair of layers)."""
    orig = O._z_stencil

    def bad(Z, nl):
        Z_nz, Z_m1, Z_m2, Z_p1 = orig(Z, nl)
        return Z_nz, Z_m1, Z_m1 + 0.5 * (Z_m2 - Z_m1), Z_p1

    monkeypatch.setattr(O, "_z_stencil", bad)
    r_int, *_ = _qr4c_check(pi_mesh, False)
    assert r_int > 1e-6, r_int
