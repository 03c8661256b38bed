"""pvl_point_velocity_noalloc against pvl_point_velocity, through the C ABI.

The no-allocation entry point runs the same kernels with the same chunking as
pvl_point_velocity, with its working buffers in a fixed per-GPU array, so its results must
be bit-identical.  Two calls in a row, the second smaller (the array keeps the first call's
contents), plus accumulation into an existing output and the size limit, which is refused
on the host before any launch.
"""

import ctypes as C

import numpy as np
import pytest

from pyvl_amd import _lib
from pyvl_amd import geometry as G
from pyvl_amd.cvl import Mesh
from pyvl_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _case(kind, seed):
    if kind == "sphere":
        pos, el = G.uv_sphere(64, 32, sigma=1e-3, seed=seed)
    else:
        pos, el, _ = G.flat_wing(16, 4, 8.0, 1.0)
    m = Mesh(pos.shape[0], el)
    rng = np.random.default_rng(seed)
    circ = rng.normal(size=m.n_surfaces)
    return m, pos, circ, rng


def _call(name, m, pos, circ, pts, out, accumulate=0):
    return getattr(_lib.load(), name)(
        m._device(), C.c_double(TOL), pos.cptr, circ.cptr, C.c_uint(pts.shape[0]), pts.cptr, out.cptr,
        C.c_int(accumulate), C.c_int(_lib.PVL_FP64))


def _both(m, pos, circ, pts, base=None):
    dpos, dcirc, dpts = DeviceArray.from_numpy(pos), DeviceArray.from_numpy(circ), DeviceArray.from_numpy(pts)
    acc = int(base is not None)
    want = DeviceArray.from_numpy(base) if acc else DeviceArray(pts.shape)
    got = DeviceArray.from_numpy(base) if acc else DeviceArray(pts.shape)
    _lib.check(_call("pvl_point_velocity", m, dpos, dcirc, dpts, want, acc), "pvl_point_velocity")
    _lib.check(_call("pvl_point_velocity_noalloc", m, dpos, dcirc, dpts, got, acc), "pvl_point_velocity_noalloc")
    return want.numpy(), got.numpy()


def test_noalloc_point_velocity_matches_allocating_entry_point():
    _lib.call("pvl_init", 0)
    # first call: 8,192-panel sphere, 30,000 points
    m, pos, circ, rng = _case("sphere", 3)
    pts = rng.normal(0.0, 1.5, (30000, 3))
    want, got = _both(m, pos, circ, pts)
    assert np.all(np.isfinite(want))
    assert np.array_equal(got, want)
    # second call, smaller: 64-panel wing, 500 points, over what the first call left behind
    m2, pos2, circ2, rng2 = _case("wing", 4)
    pts2 = rng2.normal(0.0, 2.0, (500, 3))
    want2, got2 = _both(m2, pos2, circ2, pts2)
    assert np.array_equal(got2, want2)
    # accumulation into an existing output
    base = rng2.normal(size=pts2.shape)
    want3, got3 = _both(m2, pos2, circ2, pts2, base)
    assert np.array_equal(got3, want3)


def test_noalloc_point_velocity_refuses_beyond_its_scratch():
    _lib.call("pvl_init", 0)
    m, pos, circ, _ = _case("sphere", 5)
    n_pts = 2_000_000  # one line chunk: 48 MB of partial sums, above the fixed 32 MiB
    dpos, dcirc = DeviceArray.from_numpy(pos), DeviceArray.from_numpy(circ)
    dpts, dout = DeviceArray.zeros((n_pts, 3)), DeviceArray((n_pts, 3))
    st = _call("pvl_point_velocity_noalloc", m, dpos, dcirc, dpts, dout)
    assert st == _lib.PVL_ERR_ARG
    assert b"more than its fixed" in _lib.load().pvl_last_error()
