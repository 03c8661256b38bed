"""GPU dense solves behind the reference's scipy call sites.

``run_solver`` factors and solves with ``scipy.linalg.lu_factor(A, overwrite_a=True)`` and
``scipy.linalg.lu_solve(decomp, b, overwrite_b=True)`` (python/pyvl/solver.py:223-227).
:func:`lu_factor` / :func:`lu_solve` keep those signatures and return values (``lu`` holds
L below the diagonal and U on and above it, ``piv`` the 0-based LAPACK interchanges), and
run on libpvl_hip.so's LU; the device factorisation travels with the returned tuple so a
following :func:`lu_solve` does not factor again.

:class:`LU` is the resident form (FP64 or FP32 factors in HBM), and :func:`solve_refine`
is the mixed-precision solve (FP32 LU on MFMA + FP64 iterative refinement).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray

FP64 = _lib.PVL_FP64
FP32 = _lib.PVL_FP32


PIVOT_NONE, PIVOT_PARTIAL, PIVOT_AUTO = 0, 1, 2  # include/pvl.h pvl_pivoting


class LU:
    """Device LU factorisation of an n x n matrix (row-major), FP64 or FP32 factors.

    ``pivoting``: True (default) = LAPACK's partial pivoting; "auto" = the same
    factorisation, computed without interchanges when partial pivoting would make none
    (checked at every column) and refactored with partial pivoting otherwise; False = no
    interchanges at all, meant for an FP32 factorisation that FP64 refinement
    (:func:`solve_refine`) corrects on diagonally strong systems such as the deflated
    closed-body ones."""

    def __init__(self, n: int, precision: int = FP64, pivoting=True):
        self.n = int(n)
        self.precision = int(precision)
        modes = {True: PIVOT_PARTIAL, False: PIVOT_NONE, "auto": PIVOT_AUTO}
        if pivoting not in modes:
            raise ValueError(f"pivoting must be True, False or 'auto', got {pivoting!r}")
        self.pivoting = pivoting
        h = C.c_void_p()
        _lib.call("pvl_lu_create", C.byref(h), C.c_int(self.n), C.c_int(self.precision))
        self._h = h
        if modes[pivoting] != PIVOT_PARTIAL:
            _lib.call("pvl_lu_set_pivoting", h, C.c_int(modes[pivoting]))
        self.info = 0

    def factor(self, a) -> "LU":
        """Factor ``a`` (FP64 DeviceArray or numpy array, n x n)."""
        d = a if isinstance(a, DeviceArray) else DeviceArray.from_numpy(np.ascontiguousarray(a, np.float64))
        if d.shape != (self.n, self.n) or d.dtype != np.float64:
            raise ValueError(f"expected a ({self.n}, {self.n}) float64 matrix, got {d.shape} {d.dtype}")
        info = C.c_int()
        st = _lib.load().pvl_lu_factor(self._h, d.cptr, C.byref(info))
        self.info = info.value
        if st == _lib.PVL_ERR_SINGULAR:
            # scipy warns (LinAlgWarning) and still returns the factors; keep that behaviour
            import warnings

            warnings.warn(f"Diagonal number {self.info} is exactly zero. Singular matrix.", RuntimeWarning,
                          stacklevel=2)
        else:
            _lib.check(st, "pvl_lu_factor")
        return self

    def solve(self, b, out: DeviceArray | None = None):
        """Solve A x = b for b of shape (n,) or (n, k); same container type as given."""
        host = not isinstance(b, DeviceArray)
        dt = np.float64 if self.precision == FP64 else np.float32
        if host:
            bb = np.ascontiguousarray(b, dt)
            shape = bb.shape
            d = DeviceArray.from_numpy(bb)
        else:
            if b.dtype != dt:
                raise ValueError(f"right-hand side must be {np.dtype(dt)} for this factorisation")
            shape = b.shape
            d = b if out is None else out
            if out is not None:
                _lib.call("pvl_memcpy_d2d", out.cptr, b.cptr, C.c_size_t(b.nbytes))
        if shape[0] != self.n:
            raise ValueError(f"right-hand side has {shape[0]} rows, expected {self.n}")
        nrhs = 1 if len(shape) == 1 else int(shape[1])
        for q0 in range(0, nrhs, 32):
            if nrhs > 32:
                raise ValueError("at most 32 right-hand sides per call")
            _lib.call("pvl_lu_solve", self._h, d.cptr, C.c_int(nrhs))
        return d.numpy() if host else d

    def pivots(self) -> np.ndarray:
        piv = np.empty(self.n, np.int32)
        _lib.call("pvl_lu_pivots", self._h, piv.ctypes.data_as(C.POINTER(C.c_int)))
        return piv

    def factors(self) -> np.ndarray:
        dt = np.float64 if self.precision == FP64 else np.float32
        d = DeviceArray((self.n, self.n), dt)
        _lib.call("pvl_lu_factors", self._h, d.cptr)
        return d.numpy()

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and _lib.alive():
                _lib._LIB.pvl_lu_destroy(self._h)
                self._h = None
        except Exception:
            pass


class _Decomp(tuple):
    """``(lu, piv)`` as scipy returns it, carrying the device factorisation."""

    _lu_handle: LU


def lu_factor(a, overwrite_a: bool = False, check_finite: bool = True):
    """``scipy.linalg.lu_factor`` on the GPU (FP64).  Returns ``(lu, piv)``."""
    a = np.asarray(a)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("expected square matrix")
    if check_finite and not np.all(np.isfinite(a)):
        raise ValueError("array must not contain infs or NaNs")
    h = LU(a.shape[0], FP64).factor(a)
    lu = h.factors()
    if overwrite_a and a.dtype == np.float64 and a.flags.c_contiguous and a.flags.writeable:
        a[...] = lu
        lu = a
    out = _Decomp((lu, h.pivots()))
    out._lu_handle = h
    return out


def lu_solve(lu_and_piv, b, trans: int = 0, overwrite_b: bool = False, check_finite: bool = True):
    """``scipy.linalg.lu_solve`` on the GPU (FP64; ``trans=0`` only, as run_solver uses)."""
    if trans != 0:
        raise NotImplementedError("only trans=0 is provided (the form run_solver calls)")
    b_arr = np.asarray(b)
    if check_finite and not np.all(np.isfinite(b_arr)):
        raise ValueError("array must not contain infs or NaNs")
    h = getattr(lu_and_piv, "_lu_handle", None)
    if h is None:
        lu, piv = lu_and_piv
        # re-factor from the LAPACK-format factors: rebuild A = P^T L U on the host
        lu = np.asarray(lu, np.float64)
        n = lu.shape[0]
        L = np.tril(lu, -1) + np.eye(n)
        U = np.triu(lu)
        a = L @ U
        perm = np.arange(n)
        for i, p in enumerate(piv):
            perm[i], perm[p] = perm[p], perm[i]
        pa = np.empty_like(a)
        pa[perm] = a
        h = LU(n, FP64).factor(pa)
    x = h.solve(b_arr.astype(np.float64))
    if overwrite_b and isinstance(b, np.ndarray) and b.dtype == np.float64 and b.flags.writeable:
        b[...] = x
        return b
    return x


def solve_refine(a: DeviceArray, b: DeviceArray, lu32: LU | None = None, tol: float = 1e-13, max_iter: int = 30,
                 out: DeviceArray | None = None):
    """Mixed-precision solve of A X = B: FP32 LU (MFMA) + FP64 iterative refinement.

    ``a``: (n, n) float64 DeviceArray; ``b``: (n,) or (n, k) float64 DeviceArray.  Refines
    until every column's FP64 residual has max|b - A x| <= tol * max|b|.  Returns ``(x,
    corrections, lu32, relative residuals)``; pass ``lu32`` back to reuse a factorisation.
    """
    n = a.shape[0]
    if lu32 is None:
        lu32 = LU(n, FP32).factor(a)
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    x = out if out is not None else DeviceArray(b.shape)
    iters = C.c_int()
    res = (C.c_double * nrhs)()
    _lib.call("pvl_solve_refine", lu32._h, a.cptr, b.cptr, x.cptr, C.c_int(nrhs), C.c_double(tol),
              C.c_int(max_iter), C.byref(iters), res)
    return x, iters.value, lu32, np.array(res[:])


def residual(a: DeviceArray, x: DeviceArray, b: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
    """R = B - A X in FP64 on the device."""
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    r = out if out is not None else DeviceArray(b.shape)
    _lib.call("pvl_residual", C.c_int(a.shape[0]), a.cptr, x.cptr, b.cptr, r.cptr, C.c_int(nrhs))
    return r


def deflation_constant(diag: np.ndarray) -> float:
    """mean |diag| / n for a closed body's diagonal block of size n."""
    d = np.asarray(diag, np.float64)
    return float(np.mean(np.abs(d)) / d.size)


def deflate(a: DeviceArray, r0: int, r1: int, c: float) -> None:
    """A[r0:r1, r0:r1] += c (closed-body rank-one deflation).

    A closed body's block has the uniform circulation as null vector (A 1 = 0); the shift
    gives that direction the singular value c (r1 - r0).  ``deflation_constant`` picks c so
    that this value is the block's mean |diagonal|, which keeps the condition number at
    that of A on the mean-free subspace; a shift of mean|diag| itself would make it grow
    with the panel count (5.8e4 at 4,096 ellipsoid panels against 37).
    """
    _lib.call("pvl_deflate", C.c_int(a.shape[0]), a.cptr, C.c_int(r0), C.c_int(r1), C.c_double(c))
