"""GPU LU against the reference's solver, scipy.linalg.lu_factor / lu_solve (LAPACK
dgetrf/dgetrs; python/pyvl/solver.py:223-227).

Pivot sequences must equal LAPACK's (first largest |a| per column); factors and solutions
agree to 1e-12 relative (different summation order inside the trailing updates).  The
mixed-precision path (FP32 LU + FP64 refinement) must reach 1e-10 relative to the FP64
solution.
"""

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as O
from pyvl_amd import DeviceArray
from pyvl_amd import geometry as G
from pyvl_amd import linalg as LA
from pyvl_amd.cvl import Mesh

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 63, 64, 65, 130, 300, 1000, 1025, 1500, 2100])
def test_factor_matches_lapack(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n))
    lu_ref, piv_ref = sla.lu_factor(a)
    h = LA.LU(n).factor(a)
    assert np.array_equal(h.pivots(), piv_ref)
    lu = h.factors()
    assert _rel(lu, lu_ref) < 1e-12


def test_ties_pick_first_row():
    """idamax semantics: equal |a| -> the first row."""
    a = np.array([[1.0, 2.0, 0.5, 1.0], [3.0, 1.0, 1.0, 0.0], [-3.0, 0.0, 2.0, 1.0], [2.0, 1.0, 0.0, 4.0]])
    _, piv_ref = sla.lu_factor(a)
    assert np.array_equal(LA.LU(4).factor(a).pivots(), piv_ref)


@pytest.mark.parametrize("n,nrhs", [(7, 1), (64, 3), (257, 16), (1000, 32), (1500, 1)])
def test_solve_matches_lapack(n, nrhs):
    rng = np.random.default_rng(7 + n)
    a = rng.normal(size=(n, n)) + n ** 0.5 * np.eye(n)
    b = rng.normal(size=(n, nrhs)) if nrhs > 1 else rng.normal(size=n)
    x_ref = sla.lu_solve(sla.lu_factor(a), b)
    x = LA.LU(n).factor(a).solve(b)
    assert _rel(x, x_ref) < 1e-12


def _solve_with_factors(lu, piv, b):
    """x from given LU factors (LAPACK layout and pivots), in FP64 on the host."""
    n = lu.shape[0]
    perm = np.arange(n)
    for i in range(n):
        perm[i], perm[piv[i]] = perm[piv[i]], perm[i]
    lo = np.tril(lu.astype(np.float64), -1) + np.eye(n)
    up = np.triu(lu.astype(np.float64))
    y = sla.solve_triangular(lo, b[perm], lower=True, unit_diagonal=True)
    return sla.solve_triangular(up, y, lower=False)


@pytest.mark.parametrize("prec,n,nrhs", [("fp64", 9000, 1), ("fp64", 3000, 5), ("fp64", 4097, 31),
                                         ("fp32", 1500, 13), ("fp32", 2100, 32), ("fp32", 4160, 4)])
def test_block_solve_paths(prec, n, nrhs):
    """Right-hand sides too large for one workgroup's LDS take the one-launch-per-triangle
    solve (workgroups per 64-row block, flags between blocks): compared with a host solve
    from the same factors -- FP64 to 1e-12 relative; FP32 (inverted diagonal blocks) to
    1e-5, its own rounding.  Ragged last blocks (n % 64 != 0), right-hand-side counts that
    are not multiples of 4, and the maximum count are covered."""
    rng = np.random.default_rng(101 + n + nrhs)
    a = rng.normal(size=(n, n)) + 2 * n ** 0.5 * np.eye(n)
    b = rng.normal(size=(n, nrhs))
    h = LA.LU(n, LA.FP64 if prec == "fp64" else LA.FP32).factor(a)
    dt = np.float64 if prec == "fp64" else np.float32
    x = h.solve(b.astype(dt)).astype(np.float64)
    x_ref = _solve_with_factors(h.factors(), h.pivots(), b)
    assert _rel(x, x_ref) < (1e-12 if prec == "fp64" else 1e-5)
    # a second solve on the same handle (flags carry the launch's epoch; nothing is cleared)
    x2 = h.solve(b.astype(dt)).astype(np.float64)
    assert np.array_equal(x2, x)


@pytest.mark.parametrize("prec,n", [("fp64", 1), ("fp64", 5), ("fp64", 31), ("fp64", 32), ("fp64", 33),
                                    ("fp64", 100), ("fp64", 1000), ("fp64", 1024), ("fp64", 2048),
                                    ("fp64", 4096), ("fp32", 77), ("fp32", 1024), ("fp32", 4096)])
def test_single_rhs_lookahead_solve(prec, n):
    """One right-hand side with x in LDS takes the look-ahead solve (32-row blocks; the
    reduction of block s split into blocks up to s - 2, formed while block s - 1 is being
    solved, and block s - 1): against a host solve from the same factors (FP64 1e-12; FP32
    1e-5, its own rounding), and against the multi-right-hand-side solve of the same column
    (one workgroup, 64-row blocks: another association of the same sums).  Ragged last
    blocks and single-block systems are covered."""
    rng = np.random.default_rng(300 + n)
    a = rng.normal(size=(n, n)) + 2 * n ** 0.5 * np.eye(n)
    b = rng.normal(size=n)
    h = LA.LU(n, LA.FP64 if prec == "fp64" else LA.FP32).factor(a)
    dt = np.float64 if prec == "fp64" else np.float32
    x = h.solve(b.astype(dt)).astype(np.float64)
    x_ref = _solve_with_factors(h.factors(), h.pivots(), b)
    assert _rel(x, x_ref) < (1e-12 if prec == "fp64" else 1e-5)
    x2 = h.solve(np.stack([b, b], axis=1).astype(dt)).astype(np.float64)
    assert _rel(x, x2[:, 0]) < (1e-13 if prec == "fp64" else 1e-6)
    assert np.array_equal(h.solve(b.astype(dt)).astype(np.float64), x)


def test_block_solve_after_refactor():
    """One FP32 handle factored twice (different matrices): the second solve uses the second
    factorisation's diagonal-block inverses, not the first's."""
    n, nrhs = 1500, 16
    rng = np.random.default_rng(5)
    h = LA.LU(n, LA.FP32)
    for shift in (2.0, 5.0):
        a = rng.normal(size=(n, n)) + shift * n ** 0.5 * np.eye(n)
        b = rng.normal(size=(n, nrhs))
        h.factor(a)
        x = h.solve(b.astype(np.float32)).astype(np.float64)
        x_ref = _solve_with_factors(h.factors(), h.pivots(), b)
        assert _rel(x, x_ref) < 1e-5


def test_scipy_call_site_dropin():
    """lu_factor(A, overwrite_a=True) then lu_solve(decomp, b, overwrite_b=True), as
    python/pyvl/solver.py:223-227 calls them."""
    rng = np.random.default_rng(3)
    a = rng.normal(size=(200, 200))
    b = rng.normal(size=200)
    x_ref = sla.lu_solve(sla.lu_factor(a.copy()), b.copy())
    a2, b2 = a.copy(), b.copy()
    decomp = LA.lu_factor(a2, overwrite_a=True)
    x = LA.lu_solve(decomp, b2, overwrite_b=True)
    assert x is b2
    assert _rel(x, x_ref) < 1e-12
    lu_ref, piv_ref = sla.lu_factor(a)
    assert np.array_equal(decomp[1], piv_ref)
    assert _rel(decomp[0], lu_ref) < 1e-12
    # a plain (lu, piv) tuple (e.g. from scipy) is accepted too
    x2 = LA.lu_solve((lu_ref, piv_ref), b)
    assert _rel(x2, x_ref) < 1e-10


def test_singular_reports_info():
    a = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 0.0, 1.0]])
    with pytest.warns(RuntimeWarning):
        h = LA.LU(3).factor(a)
    assert h.info > 0


@pytest.mark.parametrize("n,nrhs", [(300, 1), (1000, 4), (2100, 16)])
def test_fp32_refinement_reaches_fp64(n, nrhs):
    rng = np.random.default_rng(11 + n)
    a = rng.normal(size=(n, n)) + 2 * n ** 0.5 * np.eye(n)
    b = rng.normal(size=(n, nrhs))
    x_ref = sla.lu_solve(sla.lu_factor(a), b)
    da, db = DeviceArray.from_numpy(a), DeviceArray.from_numpy(b)
    x, iters, _, res = LA.solve_refine(da, db, tol=1e-13)
    xh = x.numpy()
    assert iters <= 6
    assert np.all(res <= 1e-13)
    r_true = np.max(np.abs(b - a @ xh), axis=0) / np.max(np.abs(b), axis=0)
    np.testing.assert_allclose(res, r_true, rtol=1e-2, atol=1e-14)  # both at the rounding floor
    for q in range(nrhs):
        assert _rel(xh[:, q], x_ref[:, q]) < 1e-10
    # FP32 factors alone are only ~1e-6 accurate
    lu32 = LA.LU(n, LA.FP32).factor(da)
    x32 = lu32.solve(b.astype(np.float32))
    assert 1e-9 < _rel(x32.astype(np.float64), x_ref) < 1e-3


def test_residual_kernel():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(333, 333))
    x = rng.normal(size=(333, 5))
    b = rng.normal(size=(333, 5))
    r = LA.residual(DeviceArray.from_numpy(a), DeviceArray.from_numpy(x), DeviceArray.from_numpy(b)).numpy()
    np.testing.assert_allclose(r, b - a @ x, rtol=1e-12, atol=1e-12)


def test_flat_wing_system_solve():
    """C1: steady flat wing, b = n . -(V_inf); GPU LU vs LAPACK on the same matrix."""
    pos, el, _ = G.flat_wing(16, 4, 8.0, 1.0)
    m = Mesh(pos.shape[0], el)
    n = m.surface_normal(pos)
    c = m.surface_average_vec3(pos)
    a = m.induction_matrix3(1e-6, pos, c, n)
    alpha = np.radians(5.0)
    vinf = np.array([np.cos(alpha), 0.0, np.sin(alpha)])
    b = np.sum(n * -vinf[None, :], axis=1)
    x_ref = sla.lu_solve(sla.lu_factor(a), b)
    x = LA.lu_solve(LA.lu_factor(a), b)
    assert _rel(x, x_ref) < 1e-13
    np.testing.assert_allclose(x, -b / (2 * np.pi), rtol=1e-12)


def test_closed_body_deflated_solve():
    """Closed sphere: A 1 ~ 0.  Deflated (A + c 1 1^T) x = b, then mean removal
    (python/pyvl/solver.py:229-233); compare with LAPACK on the same deflated matrix, and
    check the mean-free answer does not depend on c."""
    pos, el = G.uv_sphere(32, 16, sigma=1e-3, seed=0)
    m = Mesh(pos.shape[0], el)
    t = O.mesh_from_elements(pos.shape[0], el)
    n = O.surface_normal(t, pos)
    c = O.surface_center(t, pos)
    a = O.induction_matrix3(t, pos, c, n, 1e-6)
    b = np.sum(n * -np.array([1.0, 0.0, 0.0])[None, :], axis=1)
    sols = []
    for shift in (LA.deflation_constant(np.diag(a)), np.mean(np.abs(np.diag(a))), 1.0):
        da = DeviceArray.from_numpy(a)
        LA.deflate(da, 0, a.shape[0], shift)
        ad = da.numpy()
        np.testing.assert_array_equal(ad, a + shift)
        x = LA.LU(a.shape[0]).factor(da).solve(b)
        x -= x.mean()
        x_ref = sla.lu_solve(sla.lu_factor(a + shift), b)
        x_ref -= x_ref.mean()
        assert _rel(x, x_ref) < 1e-12
        sols.append(x)
    assert _rel(sols[0], sols[1]) < 1e-10
    assert _rel(sols[0], sols[2]) < 1e-10


def test_fp64_solve_backward_error_closed_body():
    """Deflated closed-body system (ellipsoid 4:1:0.5, 4,096 panels): the FP64 LU solve is
    backward stable like LAPACK's (relative residual within 10x of LAPACK's own, which is
    ~4e-13 here) and agrees with it."""
    pos, el = G.ellipsoid(64, 64, axes=(4.0, 1.0, 0.5), sigma_rel=1e-3, seed=0)
    m = Mesh(pos.shape[0], el)
    dp = DeviceArray.from_numpy(pos)
    nrm = m.surface_normal(dp)
    cps = m.surface_average_vec3(dp)
    n = m.n_surfaces
    a = DeviceArray((n, n))
    m.induction_matrix3(1e-6, dp, cps, nrm, a)
    ah = a.numpy()
    LA.deflate(a, 0, n, LA.deflation_constant(np.diag(ah)))
    ad = a.numpy()
    b = -(nrm.numpy() @ np.array([1.0, 0.05, 0.02]))
    x = LA.LU(n).factor(a).solve(b)
    x_ref = sla.lu_solve(sla.lu_factor(ad), b)
    rr = lambda v: np.max(np.abs(b - ad @ v)) / np.max(np.abs(b))  # noqa: E731
    assert rr(x) <= 10.0 * rr(x_ref) + 1e-14, (rr(x), rr(x_ref))
    assert _rel(x, x_ref) < 1e-11


def _lu_nopiv(a):
    """Unpivoted Doolittle LU (float64 numpy, tests only): packed L\\U."""
    f = np.array(a, np.float64)
    for j in range(f.shape[0] - 1):
        f[j + 1:, j] /= f[j, j]
        f[j + 1:, j + 1:] -= np.outer(f[j + 1:, j], f[j, j + 1:])
    return f


@pytest.mark.parametrize("n", [50, 64, 300, 2100])
def test_unpivoted_factor(n):
    """Without pivoting the FP64 factors equal the unpivoted LU and the pivots are the
    identity (a diagonally dominant matrix, where the unpivoted LU is stable)."""
    rng = np.random.default_rng(40 + n)
    a = rng.normal(size=(n, n)) + 2.0 * n ** 0.5 * np.eye(n)
    h = LA.LU(n, LA.FP64, pivoting=False).factor(a)
    assert np.array_equal(h.pivots(), np.arange(n))
    assert _rel(h.factors(), _lu_nopiv(a)) < 1e-12
    b = rng.normal(size=n)
    assert _rel(h.solve(b), np.linalg.solve(a, b)) < 1e-12


def test_unpivoted_fp32_refinement_closed_body():
    """FP32 LU without pivoting + FP64 refinement on the deflated closed-body system
    (ellipsoid 4:1:0.5, 4,096 panels, 4 right-hand sides): the refined answer agrees with
    LAPACK's FP64 solve to 1e-10 after a few corrections."""
    pos, el = G.ellipsoid(64, 64, axes=(4.0, 1.0, 0.5), sigma_rel=1e-3, seed=0)
    m = Mesh(pos.shape[0], el)
    dp = DeviceArray.from_numpy(pos)
    nrm = m.surface_normal(dp)
    cps = m.surface_average_vec3(dp)
    n = m.n_surfaces
    a = DeviceArray((n, n))
    m.induction_matrix3(1e-6, dp, cps, nrm, a)
    LA.deflate(a, 0, n, LA.deflation_constant(np.diag(a.numpy())))
    ad = a.numpy()
    dirs = np.array([[1.0, 0.0, 0.0], [0.99, 0.1, 0.0], [0.99, 0.0, 0.1], [0.98, 0.1, 0.1]])
    b = np.ascontiguousarray(-(nrm.numpy() @ dirs.T))
    lu32 = LA.LU(n, LA.FP32, pivoting=False).factor(a)
    x, iters, _, res = LA.solve_refine(a, DeviceArray.from_numpy(b), lu32, tol=1e-12)
    assert iters <= 6 and np.all(res <= 1e-12)
    x_ref = sla.lu_solve(sla.lu_factor(ad), b)
    for q in range(b.shape[1]):
        assert _rel(x.numpy()[:, q], x_ref[:, q]) < 1e-10


@pytest.mark.parametrize("n,dominant", [(300, True), (1000, True), (2100, True), (6000, True), (300, False), (1000, False),
                                        (2100, False)])
def test_auto_pivoting_is_partial_pivoting(n, dominant):
    """"auto" gives partial pivoting's factorisation: identity pivots and the unpivoted
    factors when no interchange is due (diagonally dominant), LAPACK's pivots and factors
    when one is (random matrix: refactored with partial pivoting)."""
    rng = np.random.default_rng(70 + n)
    a = rng.normal(size=(n, n))
    if dominant:
        a += 2.0 * n * np.eye(n)
    lu_ref, piv_ref = sla.lu_factor(a)
    h = LA.LU(n, LA.FP64, pivoting="auto").factor(a)
    assert np.array_equal(h.pivots(), piv_ref)
    assert np.array_equal(piv_ref == np.arange(n), np.full(n, dominant)) or not dominant
    assert _rel(h.factors(), lu_ref) < 1e-12


def test_auto_pivoting_interchange_below_the_diagonal_block():
    """An interchange due only in rows below the 64 x 64 diagonal block (row 100 beats the
    diagonal in column 5) is detected: the factorisation is LAPACK's."""
    n = 256
    rng = np.random.default_rng(5)
    a = rng.normal(size=(n, n)) * 0.01 + np.eye(n)
    a[100, 5] = 3.0
    lu_ref, piv_ref = sla.lu_factor(a)
    assert piv_ref[5] == 100
    h = LA.LU(n, LA.FP64, pivoting="auto").factor(a)
    assert np.array_equal(h.pivots(), piv_ref)
    assert _rel(h.factors(), lu_ref) < 1e-12


def test_unpivoted_fp32_refinement_large():
    """FP32 LU without pivoting at n = 12288 (764 workgroups in the first panel launch, dispatched over
    time: the diagonal block must not be overwritten while later workgroups still read it),
    refined in FP64 on a diagonally dominant system: converges to the FP64 LU solve."""
    n = 12288
    rng = np.random.default_rng(123)
    a = rng.standard_normal((n, n)) + 4.0 * np.sqrt(n) * np.eye(n)
    b = rng.standard_normal((n, 1))
    dA = DeviceArray.from_numpy(a)
    dB = DeviceArray.from_numpy(b)
    lu32 = LA.LU(n, LA.FP32, pivoting=False).factor(dA)
    x = DeviceArray((n, 1))
    _, iters, _, relres = LA.solve_refine(dA, dB, lu32, tol=1e-13, out=x)
    assert float(np.max(relres)) <= 1e-13 and iters <= 6
    lu64 = LA.LU(n, LA.FP64, pivoting="auto").factor(dA)
    assert np.array_equal(lu64.pivots(), np.arange(n))
    x64 = lu64.solve(b)
    assert _rel(x.numpy(), x64) < 1e-11


@pytest.mark.parametrize("n,nrhs", [(1, 1), (37, 3), (1000, 1), (2100, 20), (4096, 32)])
def test_residual_matches_numpy(n, nrhs):
    """FP64 residual R = B - A X (MFMA, K slices summed in order) against numpy."""
    rng = np.random.default_rng(300 + n)
    a = rng.standard_normal((n, n))
    x = rng.standard_normal((n, nrhs))
    b = rng.standard_normal((n, nrhs))
    r = LA.residual(DeviceArray.from_numpy(a), DeviceArray.from_numpy(x), DeviceArray.from_numpy(b)).numpy()
    want = b - a @ x
    # each entry is a sum of n products of size ~1: allow a few ulps of that scale
    assert np.max(np.abs(r.reshape(want.shape) - want)) <= 1e-13 * np.sqrt(n) * 10


_COOP_CHILD = """
import sys
import numpy as np
from pyvl_amd import linalg as LA
n, out = int(sys.argv[1]), sys.argv[2]
a = np.random.default_rng(n).normal(size=(n, n))
r = {}
for name, prec in (("fp64", LA.FP64), ("fp32", LA.FP32)):
    h = LA.LU(n, prec).factor(a)
    r[name + "_lu"], r[name + "_piv"] = h.factors(), h.pivots()
np.savez(out, **r)
print("child done", flush=True)
"""


def test_cooperative_launch_option_gives_identical_factors(tmp_path):
    """The grid-wide panel is an ordinary launch by default; PVL_COOPERATIVE=1 (read once
    per process, so set in a child process) launches it cooperatively.  n = 2,600 takes the
    64-column grid panel for every panel and the rocBLAS outer updates: factors and pivots
    of both launch kinds must be equal bit for bit, in FP64 and FP32."""
    import os
    import subprocess
    import sys
    from pathlib import Path

    n = 2600
    out = tmp_path / "coop.npz"
    root = Path(__file__).resolve().parents[1]
    env = dict(os.environ, PVL_COOPERATIVE="1")
    print(f"cooperative-launch child: n={n}", flush=True)
    p = subprocess.run([sys.executable, "-c", _COOP_CHILD, str(n), str(out)], cwd=root, env=env,
                       capture_output=True, text=True, timeout=150)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    got = np.load(out)
    a = np.random.default_rng(n).normal(size=(n, n))
    for name, prec in (("fp64", LA.FP64), ("fp32", LA.FP32)):
        h = LA.LU(n, prec).factor(a)
        assert np.array_equal(got[name + "_piv"], h.pivots()), name
        assert np.array_equal(got[name + "_lu"], h.factors()), name
    lu_ref, piv_ref = sla.lu_factor(a)
    assert np.array_equal(got["fp64_piv"], piv_ref)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_repeated_factorisation_is_bit_identical(prec):
    """n = 3,000 takes the rocBLAS updates (atomics off, no split-K): two factorisations of
    the same matrix, with another one of a different matrix in between, give the same
    factors and pivots bit for bit."""
    n = 3000
    p = LA.FP64 if prec == "fp64" else LA.FP32
    rng = np.random.default_rng(11)
    a, other = rng.normal(size=(n, n)), rng.normal(size=(n, n))
    h = LA.LU(n, p)
    lu1, piv1 = h.factor(a).factors(), h.pivots()
    h.factor(other)
    lu2, piv2 = h.factor(a).factors(), h.pivots()
    assert np.array_equal(piv1, piv2)
    assert np.array_equal(lu1, lu2)
    if prec == "fp64":
        lu_ref, piv_ref = sla.lu_factor(a)
        assert np.array_equal(piv1, piv_ref)
        assert _rel(lu1, lu_ref) < 1e-12
