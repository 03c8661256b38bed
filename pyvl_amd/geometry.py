"""Seeded synthetic geometries for the benchmark configurations.

The reference loads meshes through meshio/pyvista (python/pyvl/geometry.py:10-13); the
workloads here are generated in closed form instead, so the GPU box needs no mesh files:

- :func:`flat_wing`: rectangular or tapered flat plate, quads;
- :func:`uv_sphere` / :func:`ellipsoid`: closed UV bodies with triangle caps at the poles
  and seeded N(0, sigma) vertex jitter.

Every generator returns ``(positions (P, 3) float64, elements list-of-lists)`` plus extra
arrays where useful, in the element-list form that ``Mesh(n_points, connectivity)``
accepts.
"""

from __future__ import annotations

import numpy as np


def flat_wing(n_span: int, n_chord: int, span: float, chord: float, taper: float = 1.0):
    """Flat wing in z = 0: chord along +x (leading edge x = 0), span along y.

    Panel (i_chord, i_span) has index ``i_chord * n_span + i_span``; its loop runs
    clockwise seen from +z, so with the reference's normal convention
    (src/core/mesh.c:48-68) the normals are +z.  ``taper`` is tip chord / root chord with
    a straight trailing edge sweep-free quarter... chord scales linearly with |y|.

    Returns ``positions, elements, te_lines_nodes`` where the last is the (n_span, 2)
    array of trailing-edge node pairs (x = chord end) in span order.
    """
    ny, nx = n_span + 1, n_chord + 1
    y = np.linspace(-span / 2.0, span / 2.0, ny)
    frac = np.linspace(0.0, 1.0, nx)
    c_local = chord * (1.0 - (1.0 - taper) * np.abs(y) / (span / 2.0))
    pos = np.zeros((nx * ny, 3), np.float64)
    for ix in range(nx):
        pos[ix * ny : (ix + 1) * ny, 0] = frac[ix] * c_local
        pos[ix * ny : (ix + 1) * ny, 1] = y
    elements = []
    for ic in range(n_chord):
        for js in range(n_span):
            a = ic * ny + js  # (x_ic, y_js)
            b = a + 1  # (x_ic, y_js+1)
            c = a + ny + 1  # (x_ic+1, y_js+1)
            d = a + ny  # (x_ic+1, y_js)
            # clockwise seen from +z (x forward, y left): a -> b -> c -> d
            elements.append([a, b, c, d])
    te = np.array([[n_chord * ny + j, n_chord * ny + j + 1] for j in range(n_span)], np.uint32)
    return pos, elements, te


def _uv_body(n_az: int, n_pol: int, axes, sigma: float, seed: int):
    ax, ay, az = axes
    theta = np.linspace(0.0, np.pi, n_pol + 1)[1:-1]  # interior rings
    phi = 2.0 * np.pi * np.arange(n_az) / n_az
    n_ring = n_pol - 1
    pos = np.empty((2 + n_ring * n_az, 3), np.float64)
    pos[0] = (0.0, 0.0, az)  # north pole
    for r, th in enumerate(theta):
        s = slice(1 + r * n_az, 1 + (r + 1) * n_az)
        pos[s, 0] = ax * np.sin(th) * np.cos(phi)
        pos[s, 1] = ay * np.sin(th) * np.sin(phi)
        pos[s, 2] = az * np.cos(th)
    pos[-1] = (0.0, 0.0, -az)  # south pole
    if sigma > 0.0:
        rng = np.random.default_rng(seed)
        pos += rng.normal(0.0, sigma, size=pos.shape)

    def ring(r, k):
        return 1 + r * n_az + (k % n_az)

    south = pos.shape[0] - 1
    elements = []
    # Orientation: loops run so that the reference normal convention gives outward normals.
    for k in range(n_az):
        elements.append([0, ring(0, k + 1), ring(0, k)])
    for r in range(n_ring - 1):
        for k in range(n_az):
            elements.append([ring(r, k), ring(r, k + 1), ring(r + 1, k + 1), ring(r + 1, k)])
    for k in range(n_az):
        elements.append([ring(n_ring - 1, k), ring(n_ring - 1, k + 1), south])
    return pos, elements


def uv_sphere(n_az: int = 128, n_pol: int = 64, radius: float = 1.0, sigma: float = 1e-3, seed: int = 0):
    """UV sphere: ``n_az*(n_pol-2)`` quads plus ``2*n_az`` cap triangles (8,192 panels at 128x64)."""
    return _uv_body(n_az, n_pol, (radius, radius, radius), sigma, seed)


def ellipsoid(n_az: int = 128, n_pol: int = 128, axes=(4.0, 1.0, 0.5), sigma_rel: float = 1e-3, seed: int = 0):
    """Ellipsoid with x the long axis (semi-axes 4:1:0.5 by default; 16,384 panels at 128x128).

    The UV poles lie on the x axis: the generator is run with the long axis as its polar
    axis and then rotated, so the caps sit at the nose and tail.  Jitter sigma is
    ``sigma_rel`` times the mean semi-axis.
    """
    a, b, c = axes
    scale = float(np.mean(axes))
    pos, el = _uv_body(n_az, n_pol, (b, c, a), sigma_rel * scale, seed)
    # polar axis z -> x: (x, y, z) -> (z, x, y)
    pos = np.ascontiguousarray(pos[:, [2, 0, 1]])
    return pos, el


def cone_directions(n: int, half_angle_deg: float, seed: int) -> np.ndarray:
    """``n`` unit vectors uniform on the spherical cap of the given half angle about +x."""
    rng = np.random.default_rng(seed)
    cos_max = np.cos(np.radians(half_angle_deg))
    u = rng.uniform(cos_max, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(1.0 - u * u)
    return np.stack([u, s * np.cos(phi), s * np.sin(phi)], axis=1)
