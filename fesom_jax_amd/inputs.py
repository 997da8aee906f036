"""Seeded synthetic inputs for the FCT path (flow, thickness, tracers).

Recipe (one FCT step, float64):

* thickness at rest, linear free surface: ``hnode = hnode_new`` = interface-depth
  differences masked to wet layers, ``helem`` likewise from ``zbar``
  (fesom_jax/state.py:165-183);
* horizontal velocity from a streamfunction made of random real spherical harmonics of
  degree 1..8 (N(0,1) coefficients, rotated frame), differentiated per triangle with
  ``gradient_sca`` (``u = -dpsi/dy``, ``v = dpsi/dx``), scaled so the largest surface
  speed is ``speed`` m/s, times ``exp(Z/efold)`` with depth, zero below each
  triangle's bottom;
* ``w_e`` by discrete continuity (:func:`compute_w`, fesom_jax/ale.py:309-365);
* ``T = 28 exp(-d/800) - 1 + 0.5 n``, ``S = 35 + 0.5 n`` with ``d`` the layer
  mid-depth and ``n`` a seeded N(0,1) field smoothed by ``smooth_passes`` passes of
  triangle-area-weighted node averaging; dry lanes are 0.
"""

from __future__ import annotations

import numpy as np

from .mesh import Mesh
from .meshgen import random_sph_field


def node_xyz(mesh: Mesh) -> np.ndarray:
    lon, lat = mesh.coord_nod2D[:, 0], mesh.coord_nod2D[:, 1]
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=1)


def rest_thickness(mesh: Mesh):
    """``(hnode, helem)`` at rest (fesom_jax/state.py:165-183)."""
    z_n = mesh.zbar_3d_n
    dz_n = np.zeros((mesh.nod2D, mesh.nl))
    dz_n[:, :-1] = z_n[:, :-1] - z_n[:, 1:]
    dz = np.zeros((mesh.nl,))
    dz[:-1] = mesh.zbar[:-1] - mesh.zbar[1:]
    hnode = np.where(mesh.node_layer_mask, dz_n, 0.0)
    helem = np.where(mesh.elem_layer_mask, dz[None, :], 0.0)
    return hnode, helem


def node_edge_csr(mesh: Mesh):
    """Per node, its (edge, side) incidences in increasing ``2*edge+side`` order:
    ``(offsets[N+1], entries)`` with ``entries = 2*edge + side`` (side 0: the node is
    ``edges[e,0]``)."""
    N = mesh.nod2D
    flat = mesh.edges.reshape(-1).astype(np.int64)
    order = np.argsort(flat, kind="stable")
    counts = np.bincount(flat, minlength=N)
    off = np.zeros(N + 1, np.int64)
    off[1:] = np.cumsum(counts)
    return off, order.astype(np.int64)


def _csr_gather_sum(off, entries, vals_flat, nrows):
    """Row sums of a CSR list, each row accumulated in stored order from +0.0."""
    counts = np.diff(off)
    out = np.zeros((nrows,) + vals_flat.shape[1:])
    for j in range(int(counts.max()) if counts.size else 0):
        rows = np.nonzero(counts > j)[0]
        out[rows] = out[rows] + vals_flat[entries[off[rows] + j]]
    return out


def compute_w(mesh: Mesh, uv, helem):
    """Vertical velocity at interfaces by continuity (fesom_jax/ale.py:309-365):
    antisymmetric per-level edge divergence, bottom-up cumulative sum, divided by
    ``area``, masked to wet interfaces."""
    et = mesh.edge_tri
    el1, el2 = et[:, 0], et[:, 1]
    has1, has2 = el1 >= 0, el2 >= 0
    el1s = np.where(has1, el1, 0)
    el2s = np.where(has2, el2, 0)
    cross = mesh.edge_cross_dxdy
    U, V = uv[:, :, 0], uv[:, :, 1]
    lm = mesh.elem_layer_mask
    t1 = np.where(lm[el1s] & has1[:, None],
                  (V[el1s] * cross[:, 0:1] - U[el1s] * cross[:, 1:2]) * helem[el1s], 0.0)
    t2 = np.where(lm[el2s] & has2[:, None],
                  (V[el2s] * cross[:, 2:3] - U[el2s] * cross[:, 3:4]) * helem[el2s], 0.0)
    c = t1 + (-1.0) * t2
    vals = np.stack([c, -c], axis=1).reshape(-1, mesh.nl)
    off, ent = node_edge_csr(mesh)
    div = _csr_gather_sum(off, ent, vals, mesh.nod2D)
    w = np.cumsum(div[:, ::-1], axis=1)[:, ::-1]
    w = w / np.where(mesh.area > 0.0, mesh.area, 1.0)
    return np.where(mesh.node_iface_mask, w, 0.0)


def smooth_nodes(mesh: Mesh, f, passes: int):
    """``passes`` rounds of: triangle mean of its 3 nodes, then triangle-area-weighted
    mean over each node's triangles."""
    en = mesh.elem_nodes
    off = mesh.nod_in_elem2D_offsets.astype(np.int64)
    nie = mesh.nod_in_elem2D.astype(np.int64)
    ea = mesh.elem_area
    den = _csr_gather_sum(off, nie, ea, mesh.nod2D)
    for _ in range(passes):
        em = (f[en[:, 0]] + f[en[:, 1]] + f[en[:, 2]]) / 3.0
        num = _csr_gather_sum(off, nie, em * ea.reshape((-1,) + (1,) * (f.ndim - 1)), mesh.nod2D)
        f = num / den.reshape((-1,) + (1,) * (f.ndim - 1))
    return f


def make_velocity(mesh: Mesh, rng, *, speed=0.5, efold=1000.0, lmax=8):
    """``uv [elem2D, nl, 2]`` from a random spherical-harmonic streamfunction."""
    psi = random_sph_field(node_xyz(mesh), lmax, rng)
    gs = mesh.gradient_sca
    p = psi[mesh.elem_nodes]
    dpx = gs[:, 0] * p[:, 0] + gs[:, 1] * p[:, 1] + gs[:, 2] * p[:, 2]
    dpy = gs[:, 3] * p[:, 0] + gs[:, 4] * p[:, 1] + gs[:, 5] * p[:, 2]
    u0, v0 = -dpy, dpx
    smax = np.sqrt(u0 * u0 + v0 * v0).max()
    scale = speed / smax if smax > 0 else 0.0
    prof = np.zeros(mesh.nl)
    prof[:-1] = np.exp(mesh.Z / efold)
    uv = np.empty((mesh.elem2D, mesh.nl, 2))
    m = mesh.elem_layer_mask
    uv[:, :, 0] = np.where(m, (u0 * scale)[:, None] * prof[None, :], 0.0)
    uv[:, :, 1] = np.where(m, (v0 * scale)[:, None] * prof[None, :], 0.0)
    return uv


def make_tracer(mesh: Mesh, rng, kind: str, *, smooth_passes=3, amp=0.5):
    noise = rng.standard_normal((mesh.nod2D, mesh.nl))
    noise = smooth_nodes(mesh, noise, smooth_passes)
    d = np.zeros(mesh.nl)
    d[:-1] = -mesh.Z
    if kind == "T":
        base = 28.0 * np.exp(-d / 800.0) - 1.0
    elif kind == "S":
        base = np.full(mesh.nl, 35.0)
    else:
        raise ValueError(f"unknown tracer kind {kind!r}")
    return np.where(mesh.node_layer_mask, base[None, :] + amp * noise, 0.0)


def cfl_numbers(mesh: Mesh, uv, w, hnode, dt):
    """(horizontal, vertical) CFL estimates: ``|u| dt / sqrt(elem_area)`` and
    ``|w| dt / h`` over wet lanes."""
    sp = np.sqrt(uv[..., 0] ** 2 + uv[..., 1] ** 2)
    ch = float((sp * dt / np.sqrt(mesh.elem_area)[:, None]).max())
    h = np.where(hnode > 0, hnode, np.inf)
    cv = float((np.abs(w) * dt / h).max())
    return ch, cv


def make_case(mesh: Mesh, *, seed=0, tracers=("T", "S"), dt=2700.0, speed=0.5,
              efold=1000.0, max_cfl=0.5):
    """All FCT inputs for one step. Returns a dict with ``uv, w_e, helem, hnode,
    hnode_new, T (list), T_old (list), dt, cfl, vel_scale``."""
    rng = np.random.default_rng(seed)
    hnode, helem = rest_thickness(mesh)
    uv = make_velocity(mesh, rng, speed=speed, efold=efold)
    w = compute_w(mesh, uv, helem)
    ch, cv = cfl_numbers(mesh, uv, w, hnode, dt)
    fac = 1.0
    if max(ch, cv) > max_cfl:
        fac = max_cfl / max(ch, cv)
        uv = uv * fac
        w = compute_w(mesh, uv, helem)
        ch, cv = cfl_numbers(mesh, uv, w, hnode, dt)
    Ts = [make_tracer(mesh, rng, kind) for kind in tracers]
    return dict(uv=uv, w_e=w, helem=helem, hnode=hnode, hnode_new=hnode.copy(),
                T=Ts, T_old=[t.copy() for t in Ts], dt=dt, cfl=(ch, cv), vel_scale=fac)


# --------------------------------------------------------------------------
# the same recipe with the heavy [entity, level] arrays built on a GPU (torch), for
# meshes whose inputs are too large to build quickly on the host
# --------------------------------------------------------------------------
def _t_csr_sum(torch, off, entries, vals, nrows):
    """Row sums of a CSR list on the device, each row accumulated in stored order
    from +0.0 (gathers only: deterministic)."""
    counts = off[1:] - off[:-1]
    out = torch.zeros((nrows,) + tuple(vals.shape[1:]), dtype=vals.dtype, device=vals.device)
    for j in range(int(counts.max().item()) if counts.numel() else 0):
        rows = torch.nonzero(counts > j).squeeze(1)
        out[rows] = out[rows] + vals[entries[off[rows] + j]]
    return out


def make_case_device(mesh: Mesh, device, *, seed=0, tracers=("T", "S"), dt=2700.0,
                     speed=0.5, efold=1000.0, max_cfl=0.5, smooth_passes=3, amp=0.5):
    """make_case's recipe with the [entity, level] arrays computed by torch on
    ``device``; returns torch tensors. The random draws are numpy's (same seed), the
    arithmetic the same up to rounding; ``w_e`` is rescaled with the velocity rather
    than recomputed when the CFL cap applies."""
    import torch
    dev = torch.device(device)
    f64 = torch.float64
    rng = np.random.default_rng(seed)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    N, E, L = mesh.nod2D, mesh.elem2D, mesh.nl
    nlm = up(mesh.node_layer_mask)
    elm = up(mesh.elem_layer_mask)
    zn = up(mesh.zbar_3d_n)
    hnode = torch.zeros((N, L), dtype=f64, device=dev)
    hnode[:, :-1] = zn[:, :-1] - zn[:, 1:]
    hnode = torch.where(nlm, hnode, torch.zeros((), dtype=f64, device=dev))
    dz = np.zeros(L)
    dz[:-1] = mesh.zbar[:-1] - mesh.zbar[1:]
    helem = torch.where(elm, up(dz)[None, :].expand(E, L), torch.zeros((), dtype=f64, device=dev))
    # velocity
    psi = random_sph_field(node_xyz(mesh), 8, rng)
    gs = mesh.gradient_sca
    pp = psi[mesh.elem_nodes]
    dpx = gs[:, 0] * pp[:, 0] + gs[:, 1] * pp[:, 1] + gs[:, 2] * pp[:, 2]
    dpy = gs[:, 3] * pp[:, 0] + gs[:, 4] * pp[:, 1] + gs[:, 5] * pp[:, 2]
    u0, v0 = -dpy, dpx
    smax = np.sqrt(u0 * u0 + v0 * v0).max()
    scale = speed / smax if smax > 0 else 0.0
    prof = np.zeros(L)
    prof[:-1] = np.exp(mesh.Z / efold)
    pr = up(prof)[None, :]
    zero = torch.zeros((), dtype=f64, device=dev)
    uv = torch.empty((E, L, 2), dtype=f64, device=dev)
    uv[:, :, 0] = torch.where(elm, up(u0 * scale)[:, None] * pr, zero)
    uv[:, :, 1] = torch.where(elm, up(v0 * scale)[:, None] * pr, zero)
    # w by continuity
    et = mesh.edge_tri
    el1 = up(np.where(et[:, 0] >= 0, et[:, 0], 0).astype(np.int64))
    el2 = up(np.where(et[:, 1] >= 0, et[:, 1], 0).astype(np.int64))
    has1 = up(et[:, 0] >= 0)[:, None]
    has2 = up(et[:, 1] >= 0)[:, None]
    cr = up(mesh.edge_cross_dxdy)
    U, V = uv[:, :, 0], uv[:, :, 1]
    t1 = torch.where(elm[el1] & has1, (V[el1] * cr[:, 0:1] - U[el1] * cr[:, 1:2]) * helem[el1], zero)
    t2 = torch.where(elm[el2] & has2, (V[el2] * cr[:, 2:3] - U[el2] * cr[:, 3:4]) * helem[el2], zero)
    c = t1 + (-1.0) * t2
    del t1, t2
    vals = torch.stack([c, -c], dim=1).reshape(-1, L)
    del c
    off, ent = node_edge_csr(mesh)
    div = _t_csr_sum(torch, up(off), up(ent), vals, N)
    del vals
    w = torch.flip(torch.cumsum(torch.flip(div, [1]), 1), [1])
    area = up(mesh.area)
    w = w / torch.where(area > 0, area, torch.ones((), dtype=f64, device=dev))
    w = torch.where(up(mesh.node_iface_mask), w, zero)
    # CFL cap
    sp = torch.sqrt(uv[..., 0] ** 2 + uv[..., 1] ** 2)
    ch = float((sp * dt / up(np.sqrt(mesh.elem_area))[:, None]).max())
    hinf = torch.where(hnode > 0, hnode, torch.full((), float("inf"), dtype=f64, device=dev))
    cv = float((w.abs() * dt / hinf).max())
    fac = 1.0
    if max(ch, cv) > max_cfl:
        fac = max_cfl / max(ch, cv)
        uv = uv * fac
        w = w * fac
        ch, cv = ch * fac, cv * fac
    # tracers
    noff = up(mesh.nod_in_elem2D_offsets.astype(np.int64))
    nie = up(mesh.nod_in_elem2D.astype(np.int64))
    ea = up(mesh.elem_area)
    en = up(mesh.elem_nodes.astype(np.int64))
    den = _t_csr_sum(torch, noff, nie, ea, N)
    d = np.zeros(L)
    d[:-1] = -mesh.Z
    Ts = []
    for kind in tracers:
        f = up(rng.standard_normal((N, L)))
        for _ in range(smooth_passes):
            em = (f[en[:, 0]] + f[en[:, 1]] + f[en[:, 2]]) / 3.0
            f = _t_csr_sum(torch, noff, nie, em * ea[:, None], N) / den[:, None]
        if kind == "T":
            base = up(28.0 * np.exp(-d / 800.0) - 1.0)
        elif kind == "S":
            base = up(np.full(L, 35.0))
        else:
            raise ValueError(f"unknown tracer kind {kind!r}")
        Ts.append(torch.where(nlm, base[None, :] + amp * f, zero))
    return dict(uv=uv.contiguous(), w_e=w.contiguous(), helem=helem.contiguous(),
                hnode=hnode.contiguous(), hnode_new=hnode.clone(), T=Ts,
                T_old=[t.clone() for t in Ts], dt=dt, cfl=(ch, cv), vel_scale=fac)


def make_ice_case(mesh: Mesh, *, seed=0, speed=0.3, lmax=6):
    """Seeded sea-ice state on the surface nodes for the ice FCT
    (fesom_jax/ice_adv.py:177-194): concentration ``a_ice`` in [0, 1] with sharp ice
    edges, ice volume ``m_ice`` (m) and snow volume ``m_snow`` (m) per unit area, and a
    smooth nodal ice velocity ``(u_ice, v_ice)`` scaled to ``speed`` m/s at its maximum.
    Host numpy arrays."""
    rng = np.random.default_rng(seed)
    xyz = node_xyz(mesh)
    edge = random_sph_field(xyz, lmax, rng)
    a = np.clip(0.5 + 2.0 * edge / np.abs(edge).max(), 0.0, 1.0)
    thick = 1.0 + 0.5 * random_sph_field(xyz, 4, rng) / 3.0
    m_ice = a * np.clip(thick, 0.2, None)
    m_snow = 0.2 * m_ice * (1.0 + 0.2 * np.tanh(random_sph_field(xyz, 3, rng)))
    u = random_sph_field(xyz, lmax, rng)
    v = random_sph_field(xyz, lmax, rng)
    scale = speed / max(np.abs(u).max(), np.abs(v).max())
    return dict(a_ice=a, m_ice=m_ice, m_snow=m_snow, u_ice=u * scale, v_ice=v * scale)
