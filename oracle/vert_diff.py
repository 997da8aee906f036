"""CPU oracle (test infrastructure only): a numpy float64 restatement of the reference's
implicit vertical tracer diffusion, fesom_jax/tracer_diff.py:68-141, and of the
salinity floor that follows it in the step, fesom_jax/step.py:452-456.

Only ``tests/`` uses this module, and only as the checker; the product path
(``fesom_jax_amd``) never imports it.

Operation order follows the reference's expressions left to right, no fused
multiply-add, so a GPU kernel compiled without FP contraction reproduces it bit for bit.

Pinning (tests/test_oracle_vdiff.py): backward-Euler flux form of the solved system
(residual of the implicit equation written per interface, independently of the
tridiagonal coefficients), constant-in-z tracer unchanged and column integral
conserved (fesom_jax/tests/test_tracers.py:214-235), Kv = 0 is the identity, closed
forms for the surface increment and the shortwave divergence at Kv = 0, ``Z3d`` equal
to the static depths reproduces the static path exactly; planted mistakes fail them.
"""

from __future__ import annotations

import numpy as np

from .fct_numpy import DT_DEFAULT, _shift_down, _shift_up, mask_below_bottom, tdma

S_FLOOR = 0.5                               # fesom_jax/step.py:49


def _spacings(mesh, Z3d):
    """Layer-centre spacings (tracer_diff.py:84-92): static ``Zp = [Z, Z[-1]]`` or the
    per-node ``Z3d``; exact zeros replaced by 1."""
    if Z3d is None:
        Z = np.asarray(mesh.Z, np.float64)
        Zp = np.concatenate([Z, Z[-1:]])
        dZ_up = (_shift_down(Zp) - Zp)[None, :]
        dZ_dn = (Zp - _shift_up(Zp))[None, :]
    else:
        Z3d = np.asarray(Z3d, np.float64)
        dZ_up = _shift_down(Z3d) - Z3d
        dZ_dn = Z3d - _shift_up(Z3d)
    dZ_up = np.where(dZ_up == 0.0, 1.0, dZ_up)
    dZ_dn = np.where(dZ_dn == 0.0, 1.0, dZ_dn)
    return dZ_up, dZ_dn


def vdiff_system(mesh, Kv, hnode_new, *, dt=DT_DEFAULT, Z3d=None):
    """The tridiagonal coefficients ``(a, b, c)`` and the masks ``(valid, surf)``
    (tracer_diff.py:94-112)."""
    nl = int(mesh.nl)
    Kv = np.asarray(Kv, np.float64)
    hnn = np.asarray(hnode_new, np.float64)
    dZ_up, dZ_dn = _spacings(mesh, Z3d)
    area = np.asarray(mesh.area, np.float64)
    av = np.asarray(mesh.areasvol, np.float64)
    safe_av = np.where(av > 0.0, av, 1.0)
    a_full = -Kv * dt / dZ_up * (area / safe_av)
    c_full = -_shift_up(Kv) * dt / dZ_dn * (_shift_up(area) / safe_av)
    k = np.arange(nl)[None, :]
    nzmin = (np.asarray(mesh.ulevels_nod2D) - 1)[:, None]
    nzmax = (np.asarray(mesh.nlevels_nod2D) - 1)[:, None]
    valid = np.asarray(mesh.node_layer_mask, bool)
    surf = k == nzmin
    bot = k == (nzmax - 1)
    a = np.where(surf, 0.0, a_full)
    c = np.where(bot, 0.0, c_full)
    b = np.where(valid, -a - c + hnn, 1.0)
    a = np.where(valid, a, 0.0)
    c = np.where(valid, c, 0.0)
    return a, b, c, valid, surf


def vdiff_rhs(mesh, T, a, c, valid, surf, *, dt=DT_DEFAULT, bc_surf=None, sw_3d=None):
    """Right-hand side: the explicit operator on T, the surface increment and the
    shortwave divergence (tracer_diff.py:113-125)."""
    T = np.asarray(T, np.float64)
    rhs = np.where(valid, a * (T - _shift_down(T)) + c * (T - _shift_up(T)), 0.0)
    if bc_surf is not None:
        rhs = rhs + np.where(surf & valid, np.asarray(bc_surf, np.float64)[:, None], 0.0)
    if sw_3d is not None:
        sw = np.asarray(sw_3d, np.float64)
        area = np.asarray(mesh.area, np.float64)
        av = np.asarray(mesh.areasvol, np.float64)
        safe_av = np.where(av > 0.0, av, 1.0)
        sw_div = (sw - _shift_up(sw) * (_shift_up(area) / safe_av)) * dt
        rhs = rhs + np.where(valid, sw_div, 0.0)
    return rhs


def impl_vert_diff_one(mesh, T, Kv, hnode_new, *, dt=DT_DEFAULT, bc_surf=None, sw_3d=None,
                       Z3d=None):
    """One tracer (tracer_diff.py:68-128): solve and add the increment on wet layers."""
    a, b, c, valid, surf = vdiff_system(mesh, Kv, hnode_new, dt=dt, Z3d=Z3d)
    rhs = vdiff_rhs(mesh, T, a, c, valid, surf, dt=dt, bc_surf=bc_surf, sw_3d=sw_3d)
    dT = tdma(a, b, c, rhs)
    return np.asarray(T, np.float64) + mask_below_bottom(dT, valid)


def impl_vert_diff(mesh, T, S, Kv, hnode_new, *, dt=DT_DEFAULT, bc_T=None, bc_S=None,
                   sw_3d=None, Z3d=None):
    """Both tracers with one Kv; shortwave on T only (tracer_diff.py:131-141)."""
    return (impl_vert_diff_one(mesh, T, Kv, hnode_new, dt=dt, bc_surf=bc_T, sw_3d=sw_3d,
                               Z3d=Z3d),
            impl_vert_diff_one(mesh, S, Kv, hnode_new, dt=dt, bc_surf=bc_S, sw_3d=None,
                               Z3d=Z3d))


def salinity_floor(mesh, S, floor=S_FLOOR):
    """``where(wet, max(S, floor), S)`` (fesom_jax/step.py:456)."""
    S = np.asarray(S, np.float64)
    return np.where(np.asarray(mesh.node_layer_mask, bool), np.maximum(S, floor), S)
