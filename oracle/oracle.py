"""ORACLE -- test infrastructure only (the checker, never the product).

Python entry to the C restatement in ``pvl_oracle.c`` (``liborc.so``) and, where it was
built, to the reference's own code compiled unmodified: the C core
(``_ref/libpyvl_core_ref.so``) and the whole C extension module (``_ref/cvl*.so``).  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module.

What pins each function: line_induction, induction_matrix3, induction_matrix,
surface_center, surface_normal, mesh_from_elements and mesh_dual are bit-equal to the
reference's own C (tests/test_oracle.py) and checked against closed forms and the reference
tests' fixtures; point_velocity against induction_matrix contracted with the circulation;
line_forces bit-equal to the reference's Mesh.line_forces (tests/test_postprocess_oracle.py);
lu_factor_small against scipy's LAPACK.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORC_PATH = HERE / "liborc.so"
REF_PATH = HERE / "_ref" / "libpyvl_core_ref.so"

MASK = 0x7FFFFFFF
REV = 0x80000000
INVALID = 0x7FFFFFFF

_ORC = None
_REF = None


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def load() -> C.CDLL:
    global _ORC
    if _ORC is None:
        if not ORC_PATH.exists():
            raise RuntimeError(f"{ORC_PATH} missing: run `make -C oracle`")
        _ORC = C.CDLL(os.fspath(ORC_PATH))
        u, d, v = C.c_uint, C.c_double, C.c_void_p
        _ORC.orc_line_induction.argtypes = [u, v, v, u, v, v, d]
        _ORC.orc_induction_matrix3.argtypes = [u, v, u, v, v, v, u, v, v, v, d]
        _ORC.orc_induction_matrix.argtypes = [u, v, u, v, v, v, u, v, v, d]
        _ORC.orc_point_velocity.argtypes = [u, v, u, v, v, v, v, u, v, v, d]
        _ORC.orc_surface_center.argtypes = [v, u, v, v, v, v]
        _ORC.orc_surface_normal.argtypes = [v, u, v, v, v, v]
        _ORC.orc_mesh_from_elements.argtypes = [u, v, v, v, v, v, v]
        _ORC.orc_mesh_dual.argtypes = [u, u, v, u, v, v, v, v, v]
        _ORC.orc_lu_factor.argtypes = [u, v, v]
        _ORC.orc_lu_factor.restype = C.c_int
        _ORC.orc_line_forces.argtypes = [u, v, v, v, v, v, v]
    return _ORC


def load_ref() -> C.CDLL | None:
    """The reference's src/core compiled from its own sources, or None if not built."""
    global _REF
    if _REF is None and REF_PATH.exists():
        _REF = C.CDLL(os.fspath(REF_PATH))
    return _REF


_REF_CVL = None


def load_ref_cvl():
    """The reference's whole C extension module ``cvl`` (src/*.c + src/core/*.c compiled
    unmodified by oracle/Makefile into _ref/), imported from its file, or None if it was not
    built.  Its Mesh and ReferenceFrame are the reference's own C types."""
    global _REF_CVL
    if _REF_CVL is None:
        import importlib.util
        import sysconfig

        path = HERE / "_ref" / ("cvl" + sysconfig.get_config_var("EXT_SUFFIX"))
        if not path.exists():
            return None
        spec = importlib.util.spec_from_file_location("cvl", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _REF_CVL = mod
    return _REF_CVL


# ---------------------------------------------------------------------------------------
# Topology as plain arrays: lines (L,2) u32, offsets (N+1,) u32, entries (E,) u32.
class Topo:
    def __init__(self, n_points, lines, offsets, entries):
        self.n_points = int(n_points)
        self.lines = np.ascontiguousarray(lines, np.uint32).reshape(-1, 2)
        self.offsets = np.ascontiguousarray(offsets, np.uint32)
        self.entries = np.ascontiguousarray(entries, np.uint32)

    @property
    def n_lines(self):
        return self.lines.shape[0]

    @property
    def n_surfaces(self):
        return self.offsets.size - 1


def mesh_from_elements(n_points: int, elements) -> Topo:
    """src/core/mesh.c:175-247 restated (O(E*L) search)."""
    counts = np.array([len(e) for e in elements], np.uint32)
    flat = np.concatenate([np.asarray(e, np.uint32) for e in elements]) if elements else np.zeros(0, np.uint32)
    n = max(int(flat.size), 1)
    lines = np.zeros((n, 2), np.uint32)
    off = np.zeros(counts.size + 1, np.uint32)
    ent = np.zeros(n, np.uint32)
    nl = C.c_uint()
    load().orc_mesh_from_elements(counts.size, _p(counts), _p(flat), C.byref(nl), _p(lines), _p(off), _p(ent))
    return Topo(n_points, lines[: nl.value], off, ent[: flat.size])


def mesh_dual(t: Topo):
    """src/core/mesh.c:70-173 restated: (dual_lines (L,2), dual_offsets, dual_entries)."""
    dl = np.zeros((max(t.n_lines, 1), 2), np.uint32)
    do = np.zeros(t.n_points + 1, np.uint32)
    de = np.zeros(max(2 * t.n_lines, 1), np.uint32)
    load().orc_mesh_dual(t.n_points, t.n_lines, _p(t.lines), t.n_surfaces, _p(t.offsets), _p(t.entries), _p(dl),
                         _p(do), _p(de))
    return dl[: t.n_lines], do, de[: int(do[-1])]


def line_forces(primal: Topo, dual_lines, circ, pos, vel) -> np.ndarray:
    """Mesh.line_forces (src/meshobject.c:1350-1393) restated: (n_lines, 3)."""
    dl = np.ascontiguousarray(dual_lines, np.uint32)
    c = np.ascontiguousarray(circ, np.float64)
    p = np.ascontiguousarray(pos, np.float64)
    v = np.ascontiguousarray(vel, np.float64)
    out = np.zeros((primal.n_lines, 3), np.float64)
    load().orc_line_forces(primal.n_lines, _p(primal.lines), _p(dl), _p(c), _p(p), _p(v), _p(out))
    return out


def line_induction(t: Topo, pos, cpts, tol) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float64)
    cpts = np.ascontiguousarray(cpts, np.float64)
    out = np.empty((cpts.shape[0], t.n_lines, 3), np.float64)
    load().orc_line_induction(t.n_lines, _p(t.lines), _p(pos), cpts.shape[0], _p(cpts), _p(out), tol)
    return out


def induction_matrix3(t: Topo, pos, cpts, normals, tol) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float64)
    cpts = np.ascontiguousarray(cpts, np.float64)
    normals = np.ascontiguousarray(normals, np.float64)
    out = np.empty((cpts.shape[0], t.n_surfaces), np.float64)
    load().orc_induction_matrix3(t.n_lines, _p(t.lines), t.n_surfaces, _p(t.offsets), _p(t.entries), _p(pos),
                                 cpts.shape[0], _p(cpts), _p(normals), _p(out), tol)
    return out


def induction_matrix(t: Topo, pos, cpts, tol) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float64)
    cpts = np.ascontiguousarray(cpts, np.float64)
    out = np.empty((cpts.shape[0], t.n_surfaces, 3), np.float64)
    load().orc_induction_matrix(t.n_lines, _p(t.lines), t.n_surfaces, _p(t.offsets), _p(t.entries), _p(pos),
                                cpts.shape[0], _p(cpts), _p(out), tol)
    return out


def point_velocity(t: Topo, pos, circ, pts, tol) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float64)
    circ = np.ascontiguousarray(circ, np.float64)
    pts = np.ascontiguousarray(pts, np.float64)
    out = np.empty((pts.shape[0], 3), np.float64)
    load().orc_point_velocity(t.n_lines, _p(t.lines), t.n_surfaces, _p(t.offsets), _p(t.entries), _p(pos),
                              _p(circ), pts.shape[0], _p(pts), _p(out), tol)
    return out


def surface_center(t: Topo, pos) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float64)
    out = np.empty((t.n_surfaces, 3), np.float64)
    load().orc_surface_center(_p(t.lines), t.n_surfaces, _p(t.offsets), _p(t.entries), _p(pos), _p(out))
    return out


def surface_normal(t: Topo, pos) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float64)
    out = np.empty((t.n_surfaces, 3), np.float64)
    load().orc_surface_normal(_p(t.lines), t.n_surfaces, _p(t.offsets), _p(t.entries), _p(pos), _p(out))
    return out


def lu_factor_small(a: np.ndarray):
    """dgetf2-style unblocked LU (row-major) for tiny cross-checks."""
    a = np.array(a, np.float64, order="C")
    piv = np.zeros(a.shape[0], np.int32)
    info = load().orc_lu_factor(a.shape[0], _p(a), _p(piv))
    return a, piv, info


# ---------------------------------------------------------------------------------------
# The reference's own C core (unmodified sources): ctypes views of mesh_t / allocator_t.
class _MeshT(C.Structure):
    _fields_ = [
        ("n_points", C.c_uint),
        ("n_lines", C.c_uint),
        ("lines", C.c_void_p),
        ("n_surfaces", C.c_uint),
        ("surface_offsets", C.c_void_p),
        ("surface_lines", C.c_void_p),
    ]


_ALLOC = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t)
_FREE = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)
_REALLOC = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t)


class _AllocatorT(C.Structure):
    _fields_ = [("allocate", _ALLOC), ("deallocate", _FREE), ("reallocate", _REALLOC), ("state", C.c_void_p)]


_libc = C.CDLL(None)
_libc.malloc.restype = C.c_void_p
_libc.malloc.argtypes = [C.c_size_t]
_libc.free.argtypes = [C.c_void_p]
_libc.realloc.restype = C.c_void_p
_libc.realloc.argtypes = [C.c_void_p, C.c_size_t]
_alloc_cb = _ALLOC(lambda st, n: _libc.malloc(n))
_free_cb = _FREE(lambda st, p: _libc.free(p))
_realloc_cb = _REALLOC(lambda st, p, n: _libc.realloc(p, n))
_ALLOCATOR = _AllocatorT(_alloc_cb, _free_cb, _realloc_cb, None)


def _mesh_struct(t: Topo) -> _MeshT:
    return _MeshT(t.n_points, t.n_lines, _p(t.lines), t.n_surfaces, _p(t.offsets), _p(t.entries))


def _copy_u32(ptr, n) -> np.ndarray:
    if n == 0:
        return np.zeros(0, np.uint32)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint32)), shape=(n,)).copy()


def ref_mesh_from_elements(n_points: int, elements) -> Topo:
    ref = load_ref()
    counts = np.array([len(e) for e in elements], np.uint32)
    flat = np.concatenate([np.asarray(e, np.uint32) for e in elements])
    m = _MeshT()
    ref.mesh_from_elements.argtypes = [C.POINTER(_MeshT), C.c_uint, C.c_void_p, C.c_void_p, C.POINTER(_AllocatorT)]
    st = ref.mesh_from_elements(C.byref(m), counts.size, _p(counts), _p(flat), C.byref(_ALLOCATOR))
    if st != 0:
        raise RuntimeError("reference mesh_from_elements failed")
    lines = _copy_u32(m.lines, 2 * m.n_lines).reshape(-1, 2)
    off = _copy_u32(m.surface_offsets, m.n_surfaces + 1)
    ent = _copy_u32(m.surface_lines, int(off[-1]))
    for p in (m.lines, m.surface_offsets, m.surface_lines):
        _libc.free(p)
    return Topo(n_points, lines, off, ent)


def ref_mesh_dual(t: Topo):
    ref = load_ref()
    src = _mesh_struct(t)
    m = _MeshT()
    ref.mesh_dual_from_primal.argtypes = [C.POINTER(_MeshT), C.POINTER(_MeshT), C.POINTER(_AllocatorT)]
    if ref.mesh_dual_from_primal(C.byref(m), C.byref(src), C.byref(_ALLOCATOR)) != 0:
        raise RuntimeError("reference mesh_dual_from_primal failed")
    lines = _copy_u32(m.lines, 2 * m.n_lines).reshape(-1, 2)
    off = _copy_u32(m.surface_offsets, m.n_surfaces + 1)
    ent = _copy_u32(m.surface_lines, int(off[-1]))
    for p in (m.lines, m.surface_offsets, m.surface_lines):
        _libc.free(p)
    return lines, off, ent


def ref_line_induction(t: Topo, pos, cpts, tol) -> np.ndarray:
    ref = load_ref()
    pos = np.ascontiguousarray(pos, np.float64)
    cpts = np.ascontiguousarray(cpts, np.float64)
    out = np.empty((cpts.shape[0], t.n_lines, 3), np.float64)
    ref.compute_line_induction.argtypes = [C.c_uint, C.c_void_p, C.c_uint, C.c_void_p, C.c_uint, C.c_void_p,
                                           C.c_void_p, C.c_double]
    ref.compute_line_induction(t.n_lines, _p(t.lines), pos.shape[0], _p(pos), cpts.shape[0], _p(cpts), _p(out), tol)
    return out


def ref_induction_matrix3(t: Topo, pos, cpts, normals, tol) -> np.ndarray:
    """compute_line_induction + line_induction_to_normal_surface_induction, as
    src/meshobject.c:462-468 calls them (with the full line buffer)."""
    ref = load_ref()
    buf = ref_line_induction(t, pos, cpts, tol)
    normals = np.ascontiguousarray(normals, np.float64)
    out = np.empty((cpts.shape[0], t.n_surfaces), np.float64)
    ref.line_induction_to_normal_surface_induction.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                                               C.c_void_p, C.c_void_p, C.c_void_p]
    ref.line_induction_to_normal_surface_induction(t.n_surfaces, _p(t.offsets), _p(t.entries), t.n_lines,
                                                   cpts.shape[0], _p(normals), _p(buf), _p(out))
    return out


def ref_induction_matrix(t: Topo, pos, cpts, tol) -> np.ndarray:
    ref = load_ref()
    buf = ref_line_induction(t, pos, cpts, tol)
    out = np.empty((cpts.shape[0], t.n_surfaces, 3), np.float64)
    ref.line_induction_to_surface_induction.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                                        C.c_void_p, C.c_void_p]
    ref.line_induction_to_surface_induction(t.n_surfaces, _p(t.offsets), _p(t.entries), t.n_lines, cpts.shape[0],
                                            _p(buf), _p(out))
    return out


def _ref_surface_vec(fn_name: str, t: Topo, pos) -> np.ndarray:
    ref = load_ref()
    pos = np.ascontiguousarray(pos, np.float64)
    ms = _mesh_struct(t)

    class _R3(C.Structure):
        _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]

    fn = getattr(ref, fn_name)
    fn.argtypes = [C.c_void_p, C.POINTER(_MeshT), C.c_uint32]
    fn.restype = _R3
    out = np.empty((t.n_surfaces, 3), np.float64)
    for s in range(t.n_surfaces):
        r = fn(_p(pos), C.byref(ms), s)
        out[s] = (r.x, r.y, r.z)
    return out


def ref_surface_center(t: Topo, pos) -> np.ndarray:
    return _ref_surface_vec("surface_center", t, pos)


def ref_surface_normal(t: Topo, pos) -> np.ndarray:
    return _ref_surface_vec("surface_normal", t, pos)


def read_msh(path) -> tuple[np.ndarray, Topo]:
    """Reader for the reference's text mesh format (src/core/mesh_io.h:11-28): ids in the
    file are 1-based with the sign as orientation; ``0`` maps to INVALID_ID."""
    import re

    txt = Path(path).read_text()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", " ", txt)
    tok = txt.split()
    it = iter(tok)
    version = int(next(it))
    assert version == 0
    n_pts, n_lines, n_surf = int(next(it)), int(next(it)), int(next(it))
    pos = np.array([float(next(it)) for _ in range(3 * n_pts)], np.float64).reshape(-1, 3)

    def pack(i: int) -> int:
        v = (abs(i) - 1) & MASK
        return v | (REV if i < 0 else 0)

    lines = np.array([pack(int(next(it))) for _ in range(2 * n_lines)], np.uint32).reshape(-1, 2)
    off = [0]
    ent = []
    for _ in range(n_surf):
        n = int(next(it))
        ent.extend(pack(int(next(it))) for _ in range(n))
        off.append(len(ent))
    return pos, Topo(n_pts, lines, np.array(off, np.uint32), np.array(ent, np.uint32))
