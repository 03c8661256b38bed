"""Concurrent calls from several host threads.

The reference releases the GIL around its compute (src/meshobject.c:462-468), and ctypes
releases it around every call into libpvl_hip.so, so two Python threads can be inside the
library at once.  The library serialises them on one lock around its shared scratch,
workspaces and streams; each thread must get exactly the results of the same calls made
one after another.  Sizes alternate so that the shared scratch buffers are regrown while
the other thread's calls are in flight.
"""

import threading

import numpy as np
import pytest

from pyvl_amd import geometry as G
from pyvl_amd import linalg as LA
from pyvl_amd.cvl import Mesh

pytestmark = pytest.mark.gpu


def _case(kind, k):
    if kind == "sphere":
        pos, el = G.uv_sphere(16 + 8 * (k % 3), 12 + 4 * (k % 2), sigma=1e-3, seed=k)
    else:
        pos, el, _ = G.flat_wing(8 + 8 * (k % 3), 4 + 2 * (k % 2), 4.0, 1.0, taper=0.7)
        pos = pos + np.array([0.0, 0.0, 0.01]) * np.sin(pos[:, :1])
    m = Mesh(pos.shape[0], el)
    nrm = m.surface_normal(pos)
    cps = m.surface_average_vec3(pos)
    pts = cps + 0.05 * nrm
    return m, pos, cps, nrm, pts


def _work(m, pos, cps, nrm, pts):
    a = m.induction_matrix3(1e-6, pos, cps, nrm)
    v = m.induction_matrix(1e-6, pos, pts)
    x = LA.LU(a.shape[0]).factor(a).solve(-nrm[:, 0].copy())
    return a, v, x


def test_two_threads_match_sequential_results():
    cases = {(kind, k): _case(kind, k) for kind in ("sphere", "wing") for k in range(6)}
    want = {key: _work(*c) for key, c in cases.items()}
    got = {}
    errors = []
    start = threading.Barrier(2)

    def run(kind):
        try:
            start.wait()
            for rep in range(3):
                for k in range(6):
                    key = (kind, k)
                    got[(key, rep)] = _work(*cases[key])
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=run, args=(kind,)) for kind in ("sphere", "wing")]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for (key, rep), res in got.items():
        for r, w in zip(res, want[key]):
            assert np.array_equal(r, w), (key, rep)
    assert len(got) == 2 * 3 * 6
