"""Closed-body deflation constant (CPU): the shift c = mean|diag| / n restores the null
direction at the scale of the diagonal, so the deflated system is as well conditioned as
A on the mean-free subspace; c = mean|diag| makes the condition number grow with n."""

import numpy as np

from oracle import oracle as O
from pyvl_amd import geometry as G
from pyvl_amd.linalg import deflation_constant


def test_deflation_constant_conditioning():
    pos, el = G.ellipsoid(32, 32, axes=(4.0, 1.0, 0.5), sigma_rel=1e-3, seed=0)
    t = O.mesh_from_elements(pos.shape[0], el)
    cps, nrm = O.surface_center(t, pos), O.surface_normal(t, pos)
    a = O.induction_matrix3(t, pos, cps, nrm, 1e-6)
    n = a.shape[0]
    s = np.linalg.svd(a, compute_uv=False)
    assert s[-1] < 1e-12 * s[0]          # closed body: uniform circulation is a null vector
    cond_free = s[0] / s[-2]             # A on the mean-free subspace
    c = deflation_constant(np.diag(a))
    assert c == np.mean(np.abs(np.diag(a))) / n
    sd = np.linalg.svd(a + c, compute_uv=False)
    assert sd[0] / sd[-1] < 2.0 * cond_free
    sw = np.linalg.svd(a + c * n, compute_uv=False)
    assert sw[0] / sw[-1] > 100.0 * cond_free
