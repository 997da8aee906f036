"""Mesh bundles for the FCT path: the derivation of the 32 bundle arrays and an
icosahedral-sphere generator that emits the same bundle.

``derive_bundle`` restates, vectorised, the derivation semantics of the reference's
scripts/prepare_mesh.py:117-300 (orient clockwise, element-ascending CSR
``nod_in_elem2D``, areas accumulated in that order, element centres with the
min-anchored cyclic rule, ``gradient_sca``, edge geometry, the MFCT up/down triangle
scan). It takes coordinates already in the rotated frame. It is checked against the
reference's shipped pi bundle in tests/test_meshgen.py.

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
ALPHA, BETA, GAMMA = 50.0, 15.0, -90.0


def _trim(dx):
    dx = np.where(dx > HALF_CYC, dx - CYC, dx)
    return np.where(dx < -HALF_CYC, dx + CYC, dx)


def _rot_matrix():
    al, be, ga = ALPHA * RAD, BETA * RAD, GAMMA * RAD
    ca, sa, cb, sb, cg, sg = (np.cos(al), np.sin(al), np.cos(be),
                              np.sin(be), np.cos(ga), np.sin(ga))
    return np.array([
        cg * ca - sg * cb * sa, cg * sa + sg * cb * ca, sg * sb,
        -sg * ca - cg * cb * sa, -sg * sa + cg * cb * ca, cg * sb,
        sb * sa, -sb * ca, cb])


def r2g(rlon, rlat):
    """Rotated -> geographic (lon, lat) radians (the transpose of the Euler rotation
    with angles 50, 15, -90 degrees)."""
    M = _rot_matrix()
    xr, yr, zr = np.cos(rlat) * np.cos(rlon), np.cos(rlat) * np.sin(rlon), np.sin(rlat)
    xg = M[0] * xr + M[3] * yr + M[6] * zr
    yg = M[1] * xr + M[4] * yr + M[7] * zr
    zg = M[2] * xr + M[5] * yr + M[8] * zr
    glat = np.arcsin(np.clip(zg, -1.0, 1.0))
    glon = np.where((yg == 0.0) & (xg == 0.0), 0.0, np.arctan2(yg, xg))
    return glon, glat


def _csr_slots(offsets, flat):
    """Slot table of a CSR list: for slot j, the rows that have a j-th entry and that
    entry's value (so a loop over slots visits each row's entries in stored order)."""
    counts = np.diff(offsets)
    out = []
    for j in range(int(counts.max()) if counts.size else 0):
        rows = np.nonzero(counts > j)[0]
        out.append((rows, flat[offsets[rows] + j]))
    return out


def _edge_x_between_bc(bx, by, cx, cy, xx, xy):
    """Vectorised test whether direction x lies in the angular sector from c to b
    (scripts/prepare_mesh.py:251-265)."""
    cr = cx * cx + cy * cy
    ok = cr != 0.0
    crs = np.where(ok, cr, 1.0)
    bxp = (bx * cx + by * cy) / crs
    byp = (-bx * cy + by * cx) / crs
    xxp = (xx * cx + xy * cy) / crs
    xyp = (-xx * cy + xy * cx) / crs
    ab = np.arctan2(byp, bxp)
    axg = np.arctan2(xyp, xxp)
    hit = ((ab > 0.0) & (axg > 0.0) & (axg < ab)) | ((ab < 0.0) & (axg < 0.0) & (axg > ab)) \
        | (ab == axg) | (axg == 0.0)
    return ok & hit


def _up_dn_scan(node, xx, xy, en, rlon, rlat, nie, offsets):
    """For each query (node, direction x) the first element of ``node``'s CSR fan
    whose sector contains x, else -1 (scripts/prepare_mesh.py:268-291)."""
    res = np.full(node.shape[0], -1, np.int32)
    todo = np.ones(node.shape[0], bool)
    deg = offsets[node + 1] - offsets[node]
    for j in range(int(deg.max()) if deg.size else 0):
        q = np.nonzero(todo & (deg > j))[0]
        if q.size == 0:
            break
        nd = node[q]
        el = nie[offsets[nd] + j]
        tri = en[el]
        js = np.where(tri[:, 0] == nd, 0, np.where(tri[:, 1] == nd, 1, 2))
        jb = np.where(js == 0, 1, 0)
        jc = np.where(js == 2, 1, 2)
        nb = tri[np.arange(q.size), jb]
        nc = tri[np.arange(q.size), jc]
        bx = _trim(rlon[nb] - rlon[nd]); by = rlat[nb] - rlat[nd]
        cx = _trim(rlon[nc] - rlon[nd]); cy = rlat[nc] - rlat[nd]
        hit = _edge_x_between_bc(bx, by, cx, cy, xx[q], xy[q])
        res[q[hit]] = el[hit]
        todo[q[hit]] = False
    return res


def derive_bundle(coord_rot, geo, coast, elem_nodes_raw, edges, edge_tri, nlevels,
                  nlevels_nod2D, zbar, depth):
    """The 32 bundle arrays + meta from primitive mesh data (rotated coordinates),
    following scripts/prepare_mesh.py:117-243."""
    N = coord_rot.shape[0]
    E = elem_nodes_raw.shape[0]
    nl = zbar.shape[0]
    rlon, rlat = coord_rot[:, 0], coord_rot[:, 1]

    en = elem_nodes_raw.astype(np.int32).copy()
    ax, ay = rlon[en[:, 0]], rlat[en[:, 0]]
    bx = _trim(rlon[en[:, 1]] - ax); by = rlat[en[:, 1]] - ay
    cx = _trim(rlon[en[:, 2]] - ax); cy = rlat[en[:, 2]] - ay
    ccw = (bx * cy - by * cx) > 0.0
    en[ccw, 1], en[ccw, 2] = elem_nodes_raw[ccw, 2], elem_nodes_raw[ccw, 1]
    n0, n1, n2 = en[:, 0], en[:, 1], en[:, 2]

    pair_node = en.reshape(-1)
    pair_elem = np.repeat(np.arange(E, dtype=np.int32), 3)
    order = np.argsort(pair_node, kind="stable")
    nie = pair_elem[order].astype(np.int32)
    counts = np.bincount(pair_node, minlength=N)
    offsets = np.zeros(N + 1, np.int32)
    offsets[1:] = np.cumsum(counts)
    slots = _csr_slots(offsets.astype(np.int64), nie)

    nlevels = nlevels.astype(np.int32)
    nlevels_nod2D = nlevels_nod2D.astype(np.int32)
    ulevels = np.ones(E, np.int32)
    ulevels_nod2D = np.ones(N, np.int32)
    nmin = np.full(N, np.iinfo(np.int32).max, np.int64)
    for rows, els in slots:
        nmin[rows] = np.minimum(nmin[rows], nlevels[els])
    nlevels_nod2D_min = np.minimum(nmin, nlevels_nod2D).astype(np.int32)
    ul_max = np.ones(N, np.int32)

    lat0, lat1, lat2 = rlat[n0], rlat[n1], rlat[n2]
    lon0, lon1, lon2 = rlon[n0], rlon[n1], rlon[n2]
    ayc = np.cos((lat0 + lat1 + lat2) / 3.0)
    aX = _trim(lon1 - lon0) * ayc; aly = lat1 - lat0
    bX = _trim(lon2 - lon0) * ayc; bly = lat2 - lat0
    elem_area = 0.5 * np.abs(aX * bly - bX * aly) * (R_EARTH ** 2)

    third = elem_area / 3.0
    k = np.arange(nl)[None, :]
    area = np.zeros((N, nl))
    for rows, els in slots:
        area[rows] = area[rows] + np.where(k < (nlevels[els].astype(np.int64) - 1)[:, None],
                                           third[els][:, None], 0.0)
    areasvol = np.where(k < (nlevels_nod2D.astype(np.int64) - 1)[:, None], area, 0.0)
    ocean_area = float(areasvol[ulevels_nod2D <= 1, 0].sum())

    axk = np.stack([lon0, lon1, lon2], axis=1)
    amin = axk.min(axis=1, keepdims=True)
    axk = np.where(axk - amin >= HALF_CYC, axk - CYC, axk)
    elem_center_x = axk.sum(axis=1) / 3.0
    elem_center_y = (lat0 + lat1 + lat2) / 3.0
    elem_cos = np.cos(elem_center_y)
    metric_factor = np.tan(elem_center_y) / R_EARTH
    _, cglat = r2g(elem_center_x, elem_center_y)
    coriolis = 2.0 * OMEGA * np.sin(cglat)
    coriolis_node = 2.0 * OMEGA * np.sin(geo[:, 1])

    dX31 = _trim(lon2 - lon0) * elem_cos; dX21 = _trim(lon1 - lon0) * elem_cos
    dY31 = lat2 - lat0; dY21 = lat1 - lat0
    df = -0.5 * R_EARTH / elem_area
    gradient_sca = np.stack([(-dY31 + dY21) * df, (dY31) * df, (-dY21) * df,
                             (dX31 - dX21) * df, (-dX31) * df, (dX21) * df], axis=1)

    e0, e1 = edges[:, 0], edges[:, 1]
    edge_dxdy = np.stack([_trim(rlon[e1] - rlon[e0]), rlat[e1] - rlat[e0]], axis=1)
    eax, ebx = rlon[e0].copy(), rlon[e1].copy()
    diff = rlon[e0] - rlon[e1]
    eax = np.where(diff > HALF_CYC, eax - CYC, eax)
    ebx = np.where(diff < -HALF_CYC, ebx - CYC, ebx)
    emx = 0.5 * (eax + ebx); emy = 0.5 * (rlat[e0] + rlat[e1])
    el1, el2 = edge_tri[:, 0], edge_tri[:, 1]
    cross = np.zeros((edges.shape[0], 4))
    bxe = _trim(elem_center_x[el1] - emx); bye = elem_center_y[el1] - emy
    cross[:, 0] = bxe * elem_cos[el1] * R_EARTH; cross[:, 1] = bye * R_EARTH
    has2 = el2 >= 0
    el2s = np.where(has2, el2, 0)
    bx2 = _trim(elem_center_x[el2s] - emx); by2 = elem_center_y[el2s] - emy
    cross[:, 2] = np.where(has2, bx2 * elem_cos[el2s] * R_EARTH, 0.0)
    cross[:, 3] = np.where(has2, by2 * R_EARTH, 0.0)

    inv_pi = 1.0 / PI_FULL
    mres = 2.0 * np.sqrt(areasvol[:, 0] * inv_pi)
    for _ in range(3):
        mean3 = (mres[n0] + mres[n1] + mres[n2]) / 3.0
        vals = mean3 * elem_area
        acc = np.zeros(N); vol = np.zeros(N)
        for rows, els in slots:
            acc[rows] = acc[rows] + vals[els]
            vol[rows] = vol[rows] + elem_area[els]
        mres = np.where(vol > 0.0, acc / np.where(vol > 0.0, vol, 1.0), 0.0)

    zbar_3d_n = np.where(k < nlevels_nod2D.astype(np.int64)[:, None], zbar[None, :], 0.0)

    x0 = _trim(rlon[e1] - rlon[e0]); x1 = rlat[e1] - rlat[e0]
    up = _up_dn_scan(e0.astype(np.int64), -x0, -x1, en, rlon, rlat, nie, offsets.astype(np.int64))
    dn = _up_dn_scan(e1.astype(np.int64), x0, x1, en, rlon, rlat, nie, offsets.astype(np.int64))
    edge_up_dn_tri = np.stack([up, dn], axis=1).astype(np.int32)

    arrays = {
        "coord_nod2D": coord_rot, "geo_coord_nod2D": geo, "coast_flag": coast,
        "nlevels_nod2D": nlevels_nod2D, "nlevels_nod2D_min": nlevels_nod2D_min,
        "ulevels_nod2D": ulevels_nod2D, "ulevels_nod2D_max": ul_max, "depth": depth,
        "mesh_resolution": mres, "coriolis_node": coriolis_node, "area": area,
        "areasvol": areasvol, "zbar_3d_n": zbar_3d_n, "nod_in_elem2D_offsets": offsets,
        "elem_nodes": en, "nlevels": nlevels, "ulevels": ulevels, "elem_area": elem_area,
        "elem_cos": elem_cos, "metric_factor": metric_factor, "coriolis": coriolis,
        "elem_center_x": elem_center_x, "elem_center_y": elem_center_y,
        "gradient_sca": gradient_sca, "edges": edges.astype(np.int32),
        "edge_tri": edge_tri.astype(np.int32), "edge_dxdy": edge_dxdy,
        "edge_cross_dxdy": cross, "edge_up_dn_tri": edge_up_dn_tri,
        "zbar": zbar, "Z": 0.5 * (zbar[:-1] + zbar[1:]), "nod_in_elem2D": nie,
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
    en = faces.astype(np.int64).copy()
    ax, ay = rlon[en[:, 0]], rlat[en[:, 0]]
    bx = _trim(rlon[en[:, 1]] - ax); by = rlat[en[:, 1]] - ay
    cx = _trim(rlon[en[:, 2]] - ax); cy = rlat[en[:, 2]] - ay
    ccw = (bx * cy - by * cx) > 0.0
    en[ccw, 1], en[ccw, 2] = faces[ccw, 2], faces[ccw, 1]

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
