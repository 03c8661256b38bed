"""The pivoted LU panel's blocked schedule (pyvl_amd/csrc/lu.hip, panel_kernel) against the
column-by-column sweep it replaces, on the CPU: a sub-block's columns updated column by column
(in registers on the GPU), the columns right of it once per sub-block from the sub-block's
U12 rebuilt from the published pivot rows -- per element the same FMAs in the same order, so
factors and pivots must agree bit for bit.  The FMA is emulated as fl(fl(a b) + c) in both
schedules (what is checked is that the two schedules perform the same operations in the
same order).  Interchanges, exactly-zero pivots (no update applied for that column) and NaN
entries are covered; the GPU kernel's factors and pivots are checked against LAPACK in
tests/test_lu_gpu.py."""

import numpy as np
import pytest


def fma(a, b, c):
    return np.float64(np.float64(a * b) + c)
def better(v1, p1, v2, p2):
    return v1 > v2 or (v1 == v2 and p1 < p2)

def sweep(P):
    L = P.copy(); m, jb = L.shape; piv = []
    for j in range(jb):
        bv, bp = -1.0, None
        for r in range(j, m):
            v = abs(L[r, j])
            if v == v and better(v, r, bv, bp if bp is not None else 1 << 60):
                bv, bp = v, r
        p = j if bp is None else bp
        piv.append(p)
        prow = L[p].copy(); rowj = L[j].copy()
        if p != j:
            L[j] = prow; L[p] = rowj
        pv = prow[j]
        if pv != 0:
            inv = 1.0 / pv
            for r in range(j + 1, m):
                l = L[r, j] * inv; L[r, j] = l
                for c in range(j + 1, jb):
                    L[r, c] = fma(-l, prow[c], L[r, c])
    return L, piv

def blocked(P, KB=8, reverse=False):
    L = P.copy(); m, jb = L.shape; piv = []
    for j0 in range(0, jb, KB):
        kb = min(KB, jb - j0)
        V = L[:, j0:j0 + kb].copy()        # registers
        zm = 0
        ucur = np.zeros((kb, jb))
        for t in range(kb):
            j = j0 + t
            bv, bp = -1.0, None
            for r in range(j, m):
                v = abs(V[r, t])
                if v == v and better(v, r, bv, bp if bp is not None else 1 << 60):
                    bv, bp = v, r
            def published(r):
                row = L[r].copy(); row[j0:j0 + kb] = V[r]; return row
            p = j if bp is None else bp
            piv.append(p)
            rowj = published(j)
            prow = rowj if p == j else published(p)
            if p != j:
                out = [c for c in range(jb) if c < j0 or c >= j0 + kb]
                L[j, out] = prow[out]; L[p, out] = rowj[out]
                V[j] = prow[j0:j0 + kb]; V[p] = rowj[j0:j0 + kb]
            ucur[t] = prow
            pv = prow[j]
            if pv != 0:
                inv = 1.0 / pv
                for r in range(j + 1, m):
                    l = V[r, t] * inv; V[r, t] = l
                    for t2 in range(t + 1, kb):
                        V[r, t2] = fma(-l, prow[j0 + t2], V[r, t2])
            else:
                zm |= 1 << t
        L[:, j0:j0 + kb] = V
        for c in range(j0 + kb, jb):
            u = [None] * kb
            for k in range(kb):
                x = ucur[k, c]
                for i in range(k):
                    if not (zm >> i) & 1:
                        x = fma(-ucur[k, j0 + i], u[i], x)
                u[k] = x; ucur[k, c] = x
        for r in range(m):
            if r >= j0 + kb:
                for c in range(j0 + kb, jb):
                    x = L[r, c]
                    for i in (reversed(range(kb)) if reverse else range(kb)):
                        if not (zm >> i) & 1:
                            x = fma(-V[r, i], ucur[i, c], x)
                    L[r, c] = x
            elif r >= j0:
                L[r, j0 + kb:jb] = ucur[r - j0, j0 + kb:jb]
    return L, piv



@pytest.mark.parametrize("case", range(24))
def test_blocked_panel_schedule_equals_column_sweep(case):
    rng = np.random.default_rng(100 + case)
    m = int(rng.integers(1, 80))
    jb = int(rng.integers(1, min(m, 64) + 1))
    P = rng.standard_normal((m, jb))
    if case % 5 == 1:
        P[:, 3 % jb] = 0.0  # a zero column: exactly-zero pivot
    if case % 7 == 2:
        P[rng.integers(0, m), 0] = np.nan
    if case % 3 == 0:
        k = min(m, jb)
        P[np.arange(k), np.arange(k)] += 5.0  # few interchanges
    A, pa = sweep(P)
    B, pb = blocked(P)
    assert pa == pb
    assert np.array_equal(A, B, equal_nan=True)


def test_a_planted_reordering_is_caught():
    """Applying a sub-block's delayed updates in the opposite order changes the rounding of
    the columns right of it: the comparison above would see such a mistake."""
    rng = np.random.default_rng(7)
    P = rng.standard_normal((40, 24))
    A, _ = sweep(P)
    B, _ = blocked(P, reverse=True)
    assert not np.array_equal(A, B)
