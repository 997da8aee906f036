"""Static mesh container and `.npy`-bundle loader for the FCT path.

Field names, shapes and conventions are those of fesom_jax/mesh.py:62-129 and the
bundle format of docs/MESH_EXPORT_LAYOUT.md:34-110 in the reference:

* indices are 0-based int32; ``edge_tri`` / ``edge_up_dn_tri`` use ``-1`` for a
  missing neighbour (fesom_jax/mesh.py:13-17);
* 3-D fields are row-major ``[n_entity, nl]`` so a node's column is contiguous
  (fesom_jax/mesh.py:18-20);
* node ``n`` has layers ``ulevels_nod2D-1 <= k < nlevels_nod2D-1`` and interfaces
  ``ulevels_nod2D-1 <= k < nlevels_nod2D`` (fesom_jax/mesh.py:22-33, 135-151).

Here the arrays are host numpy arrays; the GPU plan (``fesom_jax_amd.plan``) copies
what the kernels need to device memory once per mesh.
"""

from __future__ import annotations

import dataclasses
from pathlib import Path
from typing import Mapping

import numpy as np

from .config import CYCLIC_LENGTH_RAD


def _meta(static: bool = False) -> dataclasses.Field:
    return dataclasses.field(metadata={"static": static})


@dataclasses.dataclass(frozen=True, eq=False)
class Mesh:
    """Frozen mesh + geometry. Array fields mirror fesom_jax/mesh.py:70-114."""

    # --- nodes ---
    coord_nod2D: np.ndarray          # (nod2D,2) f8 rotated (lon,lat) radians
    geo_coord_nod2D: np.ndarray      # (nod2D,2) f8 geographic (lon,lat) radians
    coast_flag: np.ndarray           # (nod2D,)  i4
    nlevels_nod2D: np.ndarray        # (nod2D,)  i4 max over cells
    nlevels_nod2D_min: np.ndarray    # (nod2D,)  i4 min over cells
    ulevels_nod2D: np.ndarray        # (nod2D,)  i4
    ulevels_nod2D_max: np.ndarray    # (nod2D,)  i4
    depth: np.ndarray                # (nod2D,)  f8
    mesh_resolution: np.ndarray      # (nod2D,)  f8
    coriolis_node: np.ndarray        # (nod2D,)  f8
    area: np.ndarray                 # (nod2D,nl) f8
    areasvol: np.ndarray             # (nod2D,nl) f8
    zbar_3d_n: np.ndarray            # (nod2D,nl) f8
    nod_in_elem2D_offsets: np.ndarray  # (nod2D+1,) i4

    # --- elements ---
    elem_nodes: np.ndarray           # (elem2D,3) i4, clockwise
    nlevels: np.ndarray              # (elem2D,) i4
    ulevels: np.ndarray              # (elem2D,) i4
    elem_area: np.ndarray            # (elem2D,) f8
    elem_cos: np.ndarray             # (elem2D,) f8
    metric_factor: np.ndarray        # (elem2D,) f8
    coriolis: np.ndarray             # (elem2D,) f8
    elem_center_x: np.ndarray        # (elem2D,) f8
    elem_center_y: np.ndarray        # (elem2D,) f8
    gradient_sca: np.ndarray         # (elem2D,6) f8

    # --- edges ---
    edges: np.ndarray                # (edge2D,2) i4
    edge_tri: np.ndarray             # (edge2D,2) i4, -1 = boundary
    edge_dxdy: np.ndarray            # (edge2D,2) f8 radians
    edge_cross_dxdy: np.ndarray      # (edge2D,4) f8 metres
    edge_up_dn_tri: np.ndarray       # (myDim_edge2D,2) i4, -1 = absent

    # --- flat / vertical ---
    nod_in_elem2D: np.ndarray        # (sum counts,) i4
    zbar: np.ndarray                 # (nl,) f8
    Z: np.ndarray                    # (nl-1,) f8

    # --- derived masks ---
    node_layer_mask: np.ndarray      # (nod2D,nl) bool
    node_iface_mask: np.ndarray      # (nod2D,nl) bool
    elem_layer_mask: np.ndarray      # (elem2D,nl) bool
    elem_iface_mask: np.ndarray      # (elem2D,nl) bool

    # --- scalar metadata ---
    nod2D: int = _meta(static=True)
    elem2D: int = _meta(static=True)
    edge2D: int = _meta(static=True)
    nl: int = _meta(static=True)
    edge2D_in: int = _meta(static=True)
    myDim_edge2D: int = _meta(static=True)
    ocean_area: float = _meta(static=True)


MESH_ARRAY_FIELDS = tuple(
    f.name for f in dataclasses.fields(Mesh)
    if not f.metadata.get("static") and not f.name.endswith("_mask"))
MESH_META_FIELDS = tuple(f.name for f in dataclasses.fields(Mesh) if f.metadata.get("static"))


def level_masks(ulevels, nlevels, nl: int):
    """``(layer_mask, iface_mask)`` ``[n, nl]`` (fesom_jax/mesh.py:135-151)."""
    ulevels = np.asarray(ulevels).reshape(-1, 1)
    nlevels = np.asarray(nlevels).reshape(-1, 1)
    k = np.arange(nl, dtype=np.int64).reshape(1, -1)
    top = k >= (ulevels - 1)
    return top & (k < (nlevels - 1)), top & (k < nlevels)


def _trim(dx, cyc=CYCLIC_LENGTH_RAD):
    half = 0.5 * cyc
    dx = np.where(dx > half, dx - cyc, dx)
    return np.where(dx < -half, dx + cyc, dx)


def check_cw_orientation(coord_nod2D, elem_nodes, *,
                         cyclic_length_rad: float = CYCLIC_LENGTH_RAD):
    """Every triangle must be clockwise in the rotated (lon, lat) plane
    (fesom_jax/mesh.py:157-201). Returns the signed cross products; raises
    ``ValueError`` if any is >= 0."""
    coord = np.asarray(coord_nod2D)
    en = np.asarray(elem_nodes)
    ax, ay = coord[en[:, 0], 0], coord[en[:, 0], 1]
    bx = _trim(coord[en[:, 1], 0] - ax, cyclic_length_rad)
    by = coord[en[:, 1], 1] - ay
    cx = _trim(coord[en[:, 2], 0] - ax, cyclic_length_rad)
    cy = coord[en[:, 2], 1] - ay
    r = bx * cy - by * cx
    bad = int((r >= 0).sum())
    if bad:
        idx = np.nonzero(r >= 0)[0][:10].tolist()
        raise ValueError(
            f"mesh orientation: {bad}/{r.size} triangle(s) are counter-clockwise or "
            f"degenerate (r >= 0); all must be clockwise. first bad elements: {idx}")
    return r


def _read_meta_txt(path: Path) -> dict:
    out = {}
    for line in path.read_text().splitlines():
        line = line.strip()
        if line:
            key, val = line.split()
            out[key] = float(val)
    return out


def _as_host(a: np.ndarray) -> np.ndarray:
    if a.dtype.kind in "iu":
        return np.ascontiguousarray(a.astype(np.int32))
    return np.ascontiguousarray(a.astype(np.float64))


def mesh_from_arrays(arrays: Mapping[str, np.ndarray], meta: Mapping[str, float],
                     *, check_orientation: bool = True) -> Mesh:
    """Build a :class:`Mesh` from a dict of the 32 bundle arrays and the meta scalars;
    the four level masks are derived (fesom_jax/mesh.py:253-274)."""
    nl = int(meta["nl"])
    arr = {k: _as_host(np.asarray(arrays[k])) for k in MESH_ARRAY_FIELDS}
    if check_orientation:
        check_cw_orientation(arr["coord_nod2D"], arr["elem_nodes"])
    nlm, nim = level_masks(arr["ulevels_nod2D"], arr["nlevels_nod2D"], nl)
    elm, eim = level_masks(arr["ulevels"], arr["nlevels"], nl)
    return Mesh(**arr, node_layer_mask=nlm, node_iface_mask=nim,
                elem_layer_mask=elm, elem_iface_mask=eim,
                nod2D=int(meta["nod2D"]), elem2D=int(meta["elem2D"]),
                edge2D=int(meta["edge2D"]), nl=nl,
                edge2D_in=int(meta["edge2D_in"]),
                myDim_edge2D=int(meta["myDim_edge2D"]),
                ocean_area=float(meta["ocean_area"]))


def load_mesh(mesh_dir, *, check_orientation: bool = True) -> Mesh:
    """Load a mesh bundle directory: one ``<field>.npy`` per array plus ``meta.txt``
    (fesom_jax/mesh.py:225-274)."""
    mesh_dir = Path(mesh_dir)
    if not mesh_dir.is_dir():
        raise FileNotFoundError(f"mesh dir not found: {mesh_dir}")
    meta = _read_meta_txt(mesh_dir / "meta.txt")
    arrays = {k: np.load(mesh_dir / f"{k}.npy", allow_pickle=False) for k in MESH_ARRAY_FIELDS}
    return mesh_from_arrays(arrays, meta, check_orientation=check_orientation)


def save_mesh(mesh: Mesh, out_dir) -> None:
    """Write ``mesh`` as a bundle readable by :func:`load_mesh` (and by the
    reference's ``load_mesh``)."""
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    for k in MESH_ARRAY_FIELDS:
        a = np.asarray(getattr(mesh, k))
        np.save(out_dir / f"{k}.npy", a.astype("<f8") if a.dtype.kind == "f" else a.astype("<i4"))
    with open(out_dir / "meta.txt", "w") as f:
        for k in ("nod2D", "elem2D", "edge2D", "nl", "edge2D_in", "myDim_edge2D"):
            f.write(f"{k} {int(getattr(mesh, k))}\n")
        f.write("npes 1\n")
        f.write(f"ocean_area {mesh.ocean_area:.17g}\n")


def wet_counts(mesh: Mesh) -> dict:
    """Wet node-layers, node-interfaces and element-layers (the units of work and
    the byte model of the FCT step)."""
    nln = mesh.nlevels_nod2D.astype(np.int64)
    uln = mesh.ulevels_nod2D.astype(np.int64)
    ne = mesh.nlevels.astype(np.int64)
    ue = mesh.ulevels.astype(np.int64)
    sum_ln = int(np.maximum(nln - uln, 0).sum())
    return {"sum_Ln": sum_ln, "sum_In": sum_ln + mesh.nod2D,
            "sum_Le": int(np.maximum(ne - ue, 0).sum())}
