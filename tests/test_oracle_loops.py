"""The numpy oracle (oracle/fct_numpy.py, following the vectorised
fesom_jax/tracer_adv.py) against the independent C loop restatement of the original C
kernel (oracle/fct_loops.c, the semantics of the reference's own loop oracle,
fesom_jax/tests/test_tracers.py:325-516). The reference holds its JAX step to that loop
oracle at 1e-11 on smooth fields and 1e-9 absolute with the limiter active
(test_tracers.py:560-596); the same bars apply here. Different code paths (masks and
segment sums against scalar loops with explicit level ranges), so a transposed index, a
wrong mask or a dropped term in either one shows up as a difference."""

import numpy as np
import pytest

from conftest import sharp_case
from oracle import fct_numpy as O
from oracle import loops as LP


def _pair(mesh, c, j=0):
    a = O.advect_one_fct(mesh, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"],
                         c["T"][j], c["T_old"][j], dt=c["dt"])[0]
    b = LP.advect_one_fct(mesh, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"],
                          c["T"][j], c["T_old"][j], dt=c["dt"])
    return a, b


def _rel(a, b, mask):
    return float(np.abs(a - b)[mask].max() / np.abs(b[mask]).max())


@pytest.fixture(scope="module")
def icosa2():
    """A 162-node sphere (10 * 4**2 + 2 before the land carve) with coast and uneven
    bottoms: the whole path on a mesh small enough to read by hand."""
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m = icosahedral_mesh(2, 24, seed=21)
    return m, make_case(m, seed=22, dt=2700.0)


def test_smooth_pi_both_tracers(pi_mesh, pi_case):
    for j in range(2):
        a, b = _pair(pi_mesh, pi_case, j)
        assert _rel(a, b, pi_mesh.node_layer_mask) <= 1e-11, j


def test_limiter_active_pi(pi_mesh, pi_case):
    c = sharp_case(pi_mesh, pi_case, 2.0)
    a, b = _pair(pi_mesh, c)
    m = pi_mesh.node_layer_mask
    assert np.abs(a - b)[m].max() < 1e-9
    lo = O.advect_one_fct(pi_mesh, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"],
                          c["T"][0], c["T_old"][0], dt=c["dt"], return_parts=True)[2]
    assert (lo["plus_fac"][m] < 1).mean() > 0.02          # the limiter binds


def test_small_icosa_whole_path(icosa2):
    m, c = icosa2
    assert m.nod2D <= 200
    for j in range(2):
        a, b = _pair(m, c, j)
        assert _rel(a, b, m.node_layer_mask) <= 1e-11
    s = sharp_case(m, c, 3.0)
    a, b = _pair(m, s)
    assert np.abs(a - b)[m.node_layer_mask].max() < 1e-9


def test_icosa_nl70_coast_and_bathymetry(icosa4_mesh, icosa4_case):
    a, b = _pair(icosa4_mesh, icosa4_case)
    assert _rel(a, b, icosa4_mesh.node_layer_mask) <= 1e-11


def test_planted_unlimited_flux_is_caught(pi_mesh, pi_case, monkeypatch):
    """An oracle that skips the limiter (all factors 1) disagrees with the loops on the
    limiter-active state."""
    c = sharp_case(pi_mesh, pi_case, 2.0)
    real = O.zalesak_limit

    def unlimited(mesh, T, LO, adf_h, adf_v, hnode_new, **kw):
        out = list(real(mesh, T, LO, adf_h, adf_v, hnode_new, **kw))
        out[0], out[1] = adf_h, adf_v
        return tuple(out)

    monkeypatch.setattr(O, "zalesak_limit", unlimited)
    a, b = _pair(pi_mesh, c)
    assert np.abs(a - b)[pi_mesh.node_layer_mask].max() > 1e-6


def test_planted_wrong_gradient_triangle_is_caught(pi_mesh, pi_case, monkeypatch):
    """Swapping the up and down triangles of the MFCT gradients changes the result."""
    real = O.fill_up_dn_grad

    def swapped(mesh, tr_xy):
        g = np.asarray(real(mesh, tr_xy))
        return g[..., [1, 0, 3, 2]]

    monkeypatch.setattr(O, "fill_up_dn_grad", swapped)
    a, b = _pair(pi_mesh, pi_case)
    assert _rel(a, b, pi_mesh.node_layer_mask) > 1e-9


def test_compute_w_matches_c_loops(pi_mesh, pi_case, icosa4_mesh, icosa4_case):
    """compute_w (fesom_jax/ale.py:309-365) against the original C kernel's loops: edge
    scatter per wet triangle level, w[k] += w[k+1] from the bottom, / area. Same
    summation orders, so bit for bit."""
    for m, c in ((pi_mesh, pi_case), (icosa4_mesh, icosa4_case)):
        a = O.compute_w(m, c["uv"], c["helem"])
        b = LP.compute_w(m, c["uv"], c["helem"])
        assert np.array_equal(a, b)
        assert np.abs(a).max() > 0


def test_planted_compute_w_sign_is_caught(pi_mesh, pi_case, monkeypatch):
    """Flipping the right triangle's sign in the oracle breaks the agreement."""
    real_scatter = O.scatter_add

    def flipped(vals, seg, num):
        return real_scatter(np.stack([vals[:, 0], vals[:, 0]], axis=1), seg, num)

    monkeypatch.setattr(O, "scatter_add", flipped)
    a = O.compute_w(pi_mesh, pi_case["uv"], pi_case["helem"])
    b = LP.compute_w(pi_mesh, pi_case["uv"], pi_case["helem"])
    assert not np.allclose(a, b, rtol=1e-6, atol=0)
