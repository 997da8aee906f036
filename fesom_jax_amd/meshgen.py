"""Mesh bundles for the FCT path: the derivation of the 32 bundle arrays from a
triangulation in the rotated frame, and an icosahedral-sphere generator that emits the
same bundle.

``derive_bundle`` produces what the reference's mesh preparation produces
(scripts/prepare_mesh.py:117-300 defines the bundle's semantics): clockwise triangles,
node -> element lists in increasing element id, level tables, areas accumulated in that
list order, element centroids with the cyclic rule, the linear-interpolant gradient
operator, edge vectors, the smoothed resolution and the MFCT up/down triangles. It is
checked against the shipped pi bundle in tests/test_meshgen.py (integers exactly, floats
to 1e-15 of each array's largest value); the sector test of the up/down triangle search
is kept in the same arithmetic as the reference's, since its decisions on a regular grid
hang on exact ties.

``icosahedral_mesh`` builds a level-``n`` icosahedral triangulation of the sphere
(``10*4**n + 2`` nodes), treats its coordinates as the rotated frame, carves land
(polar caps of the rotated frame plus a seeded smooth field, ~30% of nodes), assigns
per-element bottom levels from a seeded smooth bathymetry, orients every edge so that
``edge_tri[:,0]`` lies to its left (the convention of the pi bundle: interior edges
first, boundary edges last, ``edge_tri[:,1] = -1`` on the coast), and numbers nodes,
elements and edges in the depth-first order of the subdivision tree so that
neighbours are close in memory.
"""

from __future__ import annotations

import math

import numpy as np

from .config import OMEGA, PI, R_EARTH, RAD
from .mesh import Mesh, mesh_from_arrays

PI_FULL = 3.14159265358979323846
CYC = 2.0 * PI
HALF_CYC = 0.5 * CYC
# Euler angles (degrees) of the rotated frame about z, x, z
ALPHA, BETA, GAMMA = 50.0, 15.0, -90.0


def _wrap(dlon):
    """Longitude differences folded into [-half cycle, half cycle]."""
    dlon = np.where(dlon > HALF_CYC, dlon - CYC, dlon)
    return np.where(dlon < -HALF_CYC, dlon + CYC, dlon)


def _frame_rotation():
    """3x3 matrix taking geographic unit vectors to the rotated frame: rotations by
    ALPHA about z, BETA about the new x, GAMMA about the new z."""
    def about_z(t):
        c, s = math.cos(t), math.sin(t)
        return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])

    def about_x(t):
        c, s = math.cos(t), math.sin(t)
        return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])

    return about_z(GAMMA * RAD) @ about_x(BETA * RAD) @ about_z(ALPHA * RAD)


def r2g(rlon, rlat):
    """Rotated -> geographic (lon, lat) radians."""
    rlon, rlat = np.asarray(rlon, np.float64), np.asarray(rlat, np.float64)
    unit = np.stack([np.cos(rlat) * np.cos(rlon), np.cos(rlat) * np.sin(rlon), np.sin(rlat)])
    x, y, z = _frame_rotation().T @ unit.reshape(3, -1)
    lat = np.arcsin(np.clip(z, -1.0, 1.0))
    lon = np.where((x == 0.0) & (y == 0.0), 0.0, np.arctan2(y, x))
    return lon.reshape(rlon.shape), lat.reshape(rlat.shape)


class _Fans:
    """Node -> element lists (increasing element id per node) of a triangulation, with
    the list walked slot by slot: slot j holds the j-th element of every node that has
    one, so a vectorised loop over slots visits each node's elements in list order."""

    def __init__(self, tri: np.ndarray, N: int):
        E = tri.shape[0]
        node_of = tri.reshape(-1)
        elem_of = np.repeat(np.arange(E, dtype=np.int64), 3)
        order = np.lexsort((elem_of, node_of))
        self.elems = elem_of[order].astype(np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(np.bincount(node_of, minlength=N))]).astype(np.int32)
        sizes = np.diff(self.offsets)
        self.slots = []
        for j in range(int(sizes.max()) if N else 0):
            nodes = np.nonzero(sizes > j)[0]
            self.slots.append((nodes, self.elems[self.offsets[nodes] + j]))

    def accumulate(self, out, value_of_elem):
        """out[n] += value_of_elem(elements) over each node's list, in list order."""
        for nodes, els in self.slots:
            out[nodes] = out[nodes] + value_of_elem(els)
        return out


def _clockwise(lon, lat, tri_raw):
    """Triangles with their last two vertices swapped where they run counter-clockwise
    in the (lon, lat) plane."""
    tri = np.array(tri_raw, dtype=np.int32, copy=True)
    v0, v1, v2 = tri_raw[:, 0], tri_raw[:, 1], tri_raw[:, 2]
    turn = (_wrap(lon[v1] - lon[v0]) * (lat[v2] - lat[v0])
            - (lat[v1] - lat[v0]) * _wrap(lon[v2] - lon[v0]))
    ccw = turn > 0.0
    tri[ccw, 1], tri[ccw, 2] = tri_raw[ccw, 2], tri_raw[ccw, 1]
    return tri


def _triangle_geometry(lon, lat, tri):
    """Per triangle: centroid (lon with the cyclic rule: vertices more than half a
    cycle east of the westernmost one are moved a cycle west), the metric factor
    cos(centroid lat), the area in m^2, and the six coefficients of the gradient of the
    linear interpolant in metres (x along cos(lat) * lon, y along lat)."""
    p = [lon[tri[:, i]] for i in range(3)]
    q = [lat[tri[:, i]] for i in range(3)]
    cy = (q[0] + q[1] + q[2]) / 3.0
    cosc = np.cos(cy)
    west = np.minimum(np.minimum(p[0], p[1]), p[2])
    shifted = [np.where(pi - west >= HALF_CYC, pi - CYC, pi) for pi in p]
    cx = (shifted[0] + shifted[1] + shifted[2]) / 3.0
    # edge vectors from vertex 0 (radians, x scaled by the centroid's metric factor)
    x1, x2 = _wrap(p[1] - p[0]) * cosc, _wrap(p[2] - p[0]) * cosc
    y1, y2 = q[1] - q[0], q[2] - q[0]
    area = 0.5 * np.abs(x1 * y2 - x2 * y1) * (R_EARTH ** 2)
    # Cramer's rule for the gradient: the determinant of the clockwise metric triangle
    # is -2 * area / R; w is R over it
    w = -0.5 * R_EARTH / area
    grad = np.stack([(y1 - y2) * w, y2 * w, -y1 * w,
                     (x2 - x1) * w, -x2 * w, x1 * w], axis=1)
    return cx, cy, cosc, area, grad


def _edge_vectors(lon, lat, edges, edge_tri, cx, cy, cosc):
    """edge_dxdy (radians, from the first node to the second) and edge_cross_dxdy:
    metres from the edge midpoint to the centroid of each adjacent triangle (x scaled by
    the triangle's metric factor), zero where the second triangle is missing."""
    a, b = edges[:, 0], edges[:, 1]
    dxdy = np.stack([_wrap(lon[b] - lon[a]), lat[b] - lat[a]], axis=1)
    la, lb = lon[a], lon[b]
    # midpoint longitude: the end that is more than half a cycle east moves a cycle west
    la = np.where(la - lb > HALF_CYC, la - CYC, la)
    lb = np.where(lon[a] - lb < -HALF_CYC, lb - CYC, lb)
    mx, my = 0.5 * (la + lb), 0.5 * (lat[a] + lat[b])
    cross = np.zeros((edges.shape[0], 4))
    for side in (0, 1):
        el = edge_tri[:, side]
        has = el >= 0
        e = np.where(has, el, 0)
        cross[:, 2 * side] = np.where(has, _wrap(cx[e] - mx) * cosc[e] * R_EARTH, 0.0)
        cross[:, 2 * side + 1] = np.where(has, (cy[e] - my) * R_EARTH, 0.0)
    return dxdy, cross


def _sector_holds(bx, by, cx, cy, dx, dy):
    """Whether direction d lies in the angle from c to b (both rays included), tested
    as the reference does: angles measured in c's frame, exact ties counted in
    (scripts/prepare_mesh.py:251-265)."""
    c2 = cx * cx + cy * cy
    nz = c2 != 0.0
    c2 = np.where(nz, c2, 1.0)
    ang_b = np.arctan2((-bx * cy + by * cx) / c2, (bx * cx + by * cy) / c2)
    ang_d = np.arctan2((-dx * cy + dy * cx) / c2, (dx * cx + dy * cy) / c2)
    inside = np.where(ang_b > 0.0, (ang_d > 0.0) & (ang_d < ang_b),
                      (ang_b < 0.0) & (ang_d < 0.0) & (ang_d > ang_b))
    return nz & (inside | (ang_b == ang_d) | (ang_d == 0.0))


def _upwind_triangles(lon, lat, tri, fans, edges):
    """edge_up_dn_tri: for each end of each edge, the first triangle in the end node's
    list whose angle at that node contains the direction pointing away from the other
    end (the MFCT stencil's upwind triangle), -1 if none."""
    a = edges[:, 0].astype(np.int64)
    b = edges[:, 1].astype(np.int64)
    dx, dy = _wrap(lon[b] - lon[a]), lat[b] - lat[a]
    sizes = np.diff(fans.offsets)
    found = []
    for node, ux, uy in ((a, -dx, -dy), (b, dx, dy)):
        res = np.full(node.shape[0], -1, np.int32)
        open_q = np.arange(node.shape[0])
        for j in range(int(sizes.max()) if sizes.size else 0):
            open_q = open_q[sizes[node[open_q]] > j] if j else open_q
            if open_q.size == 0:
                break
            nd = node[open_q]
            el = fans.elems[fans.offsets[nd] + j]
            t = tri[el]
            at = np.argmax(t == nd[:, None], axis=1)
            # the two other vertices in the order (next, then the one after)
            nb = t[np.arange(t.shape[0]), np.where(at == 0, 1, 0)]
            nc = t[np.arange(t.shape[0]), np.where(at == 2, 1, 2)]
            hit = _sector_holds(_wrap(lon[nb] - lon[nd]), lat[nb] - lat[nd],
                                _wrap(lon[nc] - lon[nd]), lat[nc] - lat[nd],
                                ux[open_q], uy[open_q])
            res[open_q[hit]] = el[hit]
            open_q = open_q[~hit]
        found.append(res)
    return np.stack(found, axis=1).astype(np.int32)


def derive_bundle(coord_rot, geo, coast, elem_nodes_raw, edges, edge_tri, nlevels,
                  nlevels_nod2D, zbar, depth):
    """The 32 bundle arrays and the meta scalars from a triangulation in the rotated
    frame: node coordinates (radians), geographic coordinates, coast flags, triangles,
    edges with their adjacent triangles, per-element and per-node bottom levels,
    interface depths and node depths."""
    N, E, nl = coord_rot.shape[0], elem_nodes_raw.shape[0], zbar.shape[0]
    lon, lat = coord_rot[:, 0], coord_rot[:, 1]
    tri = _clockwise(lon, lat, elem_nodes_raw)
    fans = _Fans(tri, N)
    nlevels = nlevels.astype(np.int32)
    nlevels_nod2D = nlevels_nod2D.astype(np.int32)
    ones_e, ones_n = np.ones(E, np.int32), np.ones(N, np.int32)   # no cavities

    # deepest-common bottom of each node's triangles (capped at the node's own bottom)
    fan_min = np.minimum.reduceat(nlevels[fans.elems], fans.offsets[:-1]) if E else ones_n
    nlevels_nod2D_min = np.minimum(fan_min, nlevels_nod2D).astype(np.int32)

    cx, cy, cosc, elem_area, grad = _triangle_geometry(lon, lat, tri)

    # node areas per level: a third of each wet triangle, summed in list order
    k = np.arange(nl)[None, :]
    third = elem_area / 3.0
    wet_third = lambda els: np.where(k < (nlevels[els].astype(np.int64) - 1)[:, None],
                                     third[els][:, None], 0.0)
    area = fans.accumulate(np.zeros((N, nl)), wet_third)
    areasvol = np.where(k < (nlevels_nod2D.astype(np.int64) - 1)[:, None], area, 0.0)
    ocean_area = float(areasvol[ones_n <= 1, 0].sum())

    _, glat_c = r2g(cx, cy)
    dxdy, cross = _edge_vectors(lon, lat, edges, edge_tri, cx, cy, cosc)

    # resolution: the diameter of a disc of the node's area, smoothed three times by
    # the area-weighted mean over the node's triangles of their vertex means
    res = 2.0 * np.sqrt(areasvol[:, 0] * (1.0 / PI_FULL))
    for _ in range(3):
        tri_mean = (res[tri[:, 0]] + res[tri[:, 1]] + res[tri[:, 2]]) / 3.0
        weighted = tri_mean * elem_area
        num = fans.accumulate(np.zeros(N), lambda els: weighted[els])
        den = fans.accumulate(np.zeros(N), lambda els: elem_area[els])
        res = np.where(den > 0.0, num / np.where(den > 0.0, den, 1.0), 0.0)

    arrays = {
        "coord_nod2D": coord_rot, "geo_coord_nod2D": geo, "coast_flag": coast,
        "nlevels_nod2D": nlevels_nod2D, "nlevels_nod2D_min": nlevels_nod2D_min,
        "ulevels_nod2D": ones_n, "ulevels_nod2D_max": ones_n.copy(), "depth": depth,
        "mesh_resolution": res, "coriolis_node": 2.0 * OMEGA * np.sin(geo[:, 1]),
        "area": area, "areasvol": areasvol,
        "zbar_3d_n": np.where(k < nlevels_nod2D.astype(np.int64)[:, None], zbar[None, :], 0.0),
        "nod_in_elem2D_offsets": fans.offsets, "elem_nodes": tri, "nlevels": nlevels,
        "ulevels": ones_e, "elem_area": elem_area, "elem_cos": cosc,
        "metric_factor": np.tan(cy) / R_EARTH, "coriolis": 2.0 * OMEGA * np.sin(glat_c),
        "elem_center_x": cx, "elem_center_y": cy, "gradient_sca": grad,
        "edges": edges.astype(np.int32), "edge_tri": edge_tri.astype(np.int32),
        "edge_dxdy": dxdy, "edge_cross_dxdy": cross,
        "edge_up_dn_tri": _upwind_triangles(lon, lat, tri, fans, edges),
        "zbar": zbar, "Z": 0.5 * (zbar[:-1] + zbar[1:]), "nod_in_elem2D": fans.elems,
    }
    meta = dict(nod2D=N, elem2D=E, edge2D=int(edges.shape[0]), nl=nl,
                edge2D_in=int(np.sum((edge_tri[:, 0] >= 0) & (edge_tri[:, 1] >= 0))),
                myDim_edge2D=int(edges.shape[0]), npes=1, ocean_area=ocean_area)
    return arrays, meta


# --------------------------------------------------------------------------
# icosahedral sphere
# --------------------------------------------------------------------------
def icosphere(level: int):
    """Unit-sphere icosahedral triangulation after ``level`` midpoint subdivisions.
    Faces are kept in depth-first subdivision order (children of face f are
    4f..4f+3); nodes are then numbered by first appearance in that face order."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                  (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)],
                 dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
                  (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
                  (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
                  (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    for _ in range(level):
        nv = v.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        pairs = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        lo = pairs.min(1); hi = pairs.max(1)
        key = lo * nv + hi
        uk, inv = np.unique(key, return_inverse=True)
        ulo, uhi = uk // nv, uk % nv
        mid = v[ulo] + v[uhi]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid])
        F = f.shape[0]
        ab = nv + inv[:F]; bc = nv + inv[F:2 * F]; ca = nv + inv[2 * F:]
        child = np.stack([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                          np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], axis=1)
        f = child.reshape(-1, 3)
    flat = f.reshape(-1)
    first = np.full(v.shape[0], flat.size, np.int64)
    np.minimum.at(first, flat, np.arange(flat.size))
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order); rank[order] = np.arange(order.size)
    return v[order], rank[f]


def flip_to_degree_eight(faces: np.ndarray, count: int, rng) -> np.ndarray:
    """Raise ``count`` degree-6 nodes of a closed triangulation to degree 8. For a node c,
    flip the two edges opposite c in two opposite triangles of its ring: each quad
    (c, u, x, v) with diagonal u-v becomes (c, u, x) + (c, x, v). c gains two
    neighbours (8), the four ring nodes u, v lose one (5), the two far vertices x gain
    one (7). Every node involved has degree 6 beforehand, and the nodes of one flip
    pair and their neighbours are not used again, so the flips do not interact."""
    f = np.array(faces, dtype=np.int64, copy=True)
    N = int(f.max()) + 1
    inc = [[] for _ in range(N)]
    for t, tri in enumerate(f):
        for v in tri:
            inc[v].append(t)
    used = np.zeros(N, bool)

    def ring(c):
        return sorted(set(f[inc[c]].reshape(-1).tolist()) - {c})

    done = 0
    for c in rng.permutation(N):
        if done == count:
            break
        c = int(c)
        if used[c] or len(inc[c]) != 6:
            continue
        r = ring(c)
        # two triangles of c's ring that share no vertex: opposite sides of the hexagon
        t0 = inc[c][0]
        e0 = [v for v in f[t0] if v != c]
        t3 = next(t for t in inc[c] if not set(f[t].tolist()) & set(e0))
        plan = []
        for t in (t0, t3):
            u, v = [int(x) for x in f[t] if x != c]
            g = next(g for g in inc[u] if g != t and v in f[g])
            x = int(next(y for y in f[g] if y != u and y != v))
            plan.append((t, g, u, v, x))
        nodes = {c, *r} | {p[4] for p in plan}
        if any(used[n] or len(inc[n]) != 6 for n in nodes):
            continue
        if any(p[4] in r for p in plan):
            continue
        for t, g, u, v, x in plan:
            f[t] = (c, u, x)
            f[g] = (c, x, v)
            inc[v].remove(t); inc[u].remove(g)
            inc[x].append(t); inc[c].append(g)
        for n in list(nodes):
            used[n] = True
            used[ring(n)] = True
        done += 1
    if done != count:
        raise ValueError(f"could only raise {done} of {count} nodes to degree 8")
    return f


def _real_sph_harm_basis(xyz, lmax):
    """Real spherical harmonics of degree 1..lmax at unit vectors (columns)."""
    from scipy.special import lpmv
    z = np.clip(xyz[:, 2], -1.0, 1.0)
    phi = np.arctan2(xyz[:, 1], xyz[:, 0])
    cols = []
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            norm = math.sqrt((2 * l + 1) / (4 * math.pi)
                             * math.factorial(l - m) / math.factorial(l + m))
            P = lpmv(m, l, z) * norm
            if m == 0:
                cols.append(P)
            else:
                cols.append(math.sqrt(2.0) * P * np.cos(m * phi))
                cols.append(math.sqrt(2.0) * P * np.sin(m * phi))
    return cols


def random_sph_field(xyz, lmax, rng, *, decay: float = 0.0):
    """Seeded random combination of real spherical harmonics of degree 1..lmax with
    N(0,1) coefficients scaled by ``l**-decay``."""
    out = np.zeros(xyz.shape[0])
    idx = 0
    basis = _real_sph_harm_basis(xyz, lmax)
    coeffs = rng.standard_normal(len(basis))
    for l in range(1, lmax + 1):
        for _ in range(2 * l + 1):
            out += coeffs[idx] * basis[idx] * (l ** -decay)
            idx += 1
    return out


def default_zbar(nl: int) -> np.ndarray:
    """Interface depths (m, <= 0) of an ``nl``-interface grid: layers thickening
    from 10 m at the surface to ~250 m, total ~6000 m."""
    nlay = nl - 1
    s = np.linspace(0.0, 1.0, nlay)
    dz = 10.0 + 240.0 * s ** 1.5
    dz *= 6000.0 / dz.sum()
    z = np.concatenate([[0.0], -np.cumsum(dz)])
    return np.round(z, 6)


def icosahedral_mesh(level: int, nl: int = 70, *, seed: int = 0,
                     land_fraction: float = 0.3, polar_cap_deg: float = 75.0,
                     min_levels: int = 10, zbar: np.ndarray | None = None,
                     degree8: int = 0) -> Mesh:
    """Ocean-only icosahedral mesh bundle (see module docstring). ``degree8`` > 0 flips
    edges so that that many nodes have 8 neighbours (``flip_to_degree_eight``), as
    unstructured ocean meshes have; 0 leaves every node at 5 or 6."""
    rng = np.random.default_rng(seed)
    xyz, faces = icosphere(level)
    if degree8:
        faces = flip_to_degree_eight(faces, degree8, np.random.default_rng(seed + 7919))
    rlat = np.arcsin(np.clip(xyz[:, 2], -1.0, 1.0))
    rlon = np.arctan2(xyz[:, 1], xyz[:, 0])

    landf = random_sph_field(xyz, 6, rng, decay=1.0)
    cap = np.abs(rlat) > polar_cap_deg * RAD
    q = np.quantile(landf[~cap], 1.0 - max(land_fraction - cap.mean(), 0.0) / max(1e-12, (~cap).mean()))
    land = cap | (landf > q)
    ocean_face = ~land[faces].any(axis=1)
    if not ocean_face.any():
        raise ValueError(f"no ocean triangle left at level {level}, land fraction "
                         f"{land_fraction}, polar caps {polar_cap_deg} deg")
    faces = faces[ocean_face]
    used = np.zeros(xyz.shape[0], bool); used[faces.reshape(-1)] = True
    remap = np.full(xyz.shape[0], -1, np.int64); remap[used] = np.arange(used.sum())
    faces = remap[faces]
    xyz = xyz[used]; rlat = rlat[used]; rlon = rlon[used]
    N = xyz.shape[0]
    E = faces.shape[0]

    coord = np.stack([rlon, rlat], axis=1)
    glon, glat = r2g(rlon, rlat)
    geo = np.stack([glon, glat], axis=1)

    # orient clockwise in the rotated plane (the derivation would swap anyway; doing
    # it here lets the edge orientation below rely on it)
    en = _clockwise(rlon, rlat, faces).astype(np.int64)

    # half-edges p->q of a clockwise triangle have the triangle on their right
    hp = np.concatenate([en[:, 0], en[:, 1], en[:, 2]])
    hq = np.concatenate([en[:, 1], en[:, 2], en[:, 0]])
    he = np.concatenate([np.arange(E)] * 3)
    hpos = np.concatenate([np.arange(E) * 3, np.arange(E) * 3 + 1, np.arange(E) * 3 + 2])
    lo = np.minimum(hp, hq); hi = np.maximum(hp, hq)
    key = lo * N + hi
    order = np.lexsort((hpos, key))
    key_s = key[order]
    start = np.r_[True, key_s[1:] != key_s[:-1]]
    gid = np.cumsum(start) - 1
    nE = int(gid[-1]) + 1
    cnt = np.bincount(gid, minlength=nE)
    first_pos = hpos[order][start]
    # for each undirected edge: the half-edge(s) and their elements
    a_idx = order[start]
    b_idx = np.full(nE, -1, np.int64)
    two = cnt == 2
    second = np.nonzero(~start)[0]
    b_idx[gid[second]] = order[second]
    interior = two
    e_n1 = np.empty(nE, np.int64); e_n2 = np.empty(nE, np.int64)
    e_t1 = np.empty(nE, np.int64); e_t2 = np.full(nE, -1, np.int64)
    # interior: orient n1 < n2; the element whose half-edge is n1->n2 is on the right (el2)
    ia = a_idx[interior]; ib = b_idx[interior]
    n1 = np.minimum(hp[ia], hq[ia]); n2 = np.maximum(hp[ia], hq[ia])
    a_is_fwd = hp[ia] == n1
    e_n1[interior] = n1; e_n2[interior] = n2
    e_t2[interior] = np.where(a_is_fwd, he[ia], he[ib])
    e_t1[interior] = np.where(a_is_fwd, he[ib], he[ia])
    # boundary: the single element must be on the left => n1->n2 = q->p
    bnd = ~interior
    ia = a_idx[bnd]
    e_n1[bnd] = hq[ia]; e_n2[bnd] = hp[ia]; e_t1[bnd] = he[ia]
    # numbering: interior edges first, then boundary edges, each in first-appearance order
    eorder = np.lexsort((first_pos, bnd.astype(np.int64)))
    edges = np.stack([e_n1[eorder], e_n2[eorder]], axis=1).astype(np.int32)
    edge_tri = np.stack([e_t1[eorder], e_t2[eorder]], axis=1).astype(np.int32)

    coast = np.zeros(N, np.int32)
    coast[edges[edge_tri[:, 1] < 0].reshape(-1)] = 1

    if zbar is None:
        zbar = default_zbar(nl)
    zbar = np.asarray(zbar, np.float64)
    nl = zbar.shape[0]
    bath = random_sph_field(xyz, 5, rng, decay=1.0)
    bath = (bath - bath.min()) / max(bath.max() - bath.min(), 1e-300)
    node_lev = np.clip(np.round(min_levels + bath * (nl - min_levels)), min_levels, nl).astype(np.int32)
    nlevels = node_lev[en].min(axis=1).astype(np.int32)
    nln = np.zeros(N, np.int32)
    np.maximum.at(nln, en.reshape(-1), np.repeat(nlevels, 3))
    depth = -zbar[node_lev - 1].astype(np.float64)

    arrays, meta = derive_bundle(coord, geo, coast, en.astype(np.int32), edges, edge_tri,
                                 nlevels, nln, zbar, depth)
    return mesh_from_arrays(arrays, meta)


def icosa_counts(level: int) -> tuple[int, int, int]:
    """(nodes, triangles, edges) of the full sphere at subdivision ``level``."""
    n = 10 * 4 ** level + 2
    return n, 20 * 4 ** level, 30 * 4 ** level


def shuffled(mesh: Mesh, seed: int) -> Mesh:
    """The same mesh with nodes, triangles and edges renumbered by seeded random
    permutations (edge orientation and triangle vertex order kept), re-derived through
    ``derive_bundle``. A caller's numbering carries no locality in general; this is the
    test case for the plan's own locality order."""
    rng = np.random.default_rng(seed)
    N, E, D = mesh.nod2D, mesh.elem2D, mesh.edge2D
    pn, pe, pd = rng.permutation(N), rng.permutation(E), rng.permutation(D)
    new_node = np.empty(N, np.int64); new_node[pn] = np.arange(N)
    new_elem = np.empty(E, np.int64); new_elem[pe] = np.arange(E)
    tri = new_node[mesh.elem_nodes[pe]].astype(np.int32)
    edges = new_node[mesh.edges[pd]].astype(np.int32)
    et = mesh.edge_tri[pd]
    edge_tri = np.where(et >= 0, new_elem[np.maximum(et, 0)], -1).astype(np.int32)
    arrays, meta = derive_bundle(mesh.coord_nod2D[pn], mesh.geo_coord_nod2D[pn],
                                 mesh.coast_flag[pn], tri, edges, edge_tri,
                                 mesh.nlevels[pe], mesh.nlevels_nod2D[pn], np.asarray(mesh.zbar),
                                 mesh.depth[pn])
    return mesh_from_arrays(arrays, meta)
