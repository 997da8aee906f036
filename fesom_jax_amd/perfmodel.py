"""Work and byte model of one FCT step (the units of the benchmark metric).

Unit of work: one node-level update = one tracer on one wet node-layer advanced one
step; per step ``U = K * sum_Ln`` with ``sum_Ln = sum(nlevels_nod2D - ulevels_nod2D)``.

Algorithmic (compulsory) HBM bytes of one step: every input read once and every
output written once, wet entries only, float64 fields, int32 topology:

    B = 8 * [ 3K*sum_Ln          T and T_old read, T_new written (T_old_new aliases T)
            + 3*sum_Le           uv (2 components) and helem
            + 2*sum_In           w_e and area (interfaces: sum_In = sum_Ln + N)
            + 3*sum_Ln ]         hnode, hnode_new, areasvol
      + 72*D + 80*E + 16*N + 4*(3E + N + 1) + 16*L

72 B per edge: edges 8 + edge_tri 8 + edge_up_dn_tri 8 + edge_cross_dxdy 32 +
edge_dxdy 16; 80 B per element: elem_nodes 12 + gradient_sca 48 + elem_area 8 +
elem_cos 8 + nlevels 4; 16 B per node of level counts; the node->element CSR list and
offsets; Z and zbar. Dry-lane copies, intermediates and the optional w_i / Z3d /
zbar3 fields are not counted.
"""

from __future__ import annotations

from .mesh import wet_counts

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s


def updates_per_step(mesh, ntracers: int) -> int:
    return int(ntracers) * wet_counts(mesh)["sum_Ln"]


def algorithmic_bytes(mesh, ntracers: int) -> int:
    w = wet_counts(mesh)
    K = int(ntracers)
    N, E, D, L = int(mesh.nod2D), int(mesh.elem2D), int(mesh.edge2D), int(mesh.nl)
    return (8 * (3 * K * w["sum_Ln"] + 3 * w["sum_Le"] + 2 * w["sum_In"] + 3 * w["sum_Ln"])
            + 72 * D + 80 * E + 16 * N + 4 * (3 * E + N + 1) + 16 * L)


def adjoint_bytes(mesh, ntracers: int) -> int:
    """Compulsory bytes of one step's backward (``fct_backward``): the step's inputs
    read once (T, T_old, flow, thicknesses, mesh), the cotangent g_T_new read, and
    g_T, g_T_old, g_uv, g_w_e written once, wet entries only:

        8 * [5K*sum_Ln + 5*sum_Le + 3*sum_In + 3*sum_Ln] + topology (as above).

    The recomputed forward intermediates are not compulsory and are not counted."""
    w = wet_counts(mesh)
    K = int(ntracers)
    N, E, D, L = int(mesh.nod2D), int(mesh.elem2D), int(mesh.edge2D), int(mesh.nl)
    return (8 * (5 * K * w["sum_Ln"] + 5 * w["sum_Le"] + 3 * w["sum_In"] + 3 * w["sum_Ln"])
            + 72 * D + 80 * E + 16 * N + 4 * (3 * E + N + 1) + 16 * L)
