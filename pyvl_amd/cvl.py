"""Drop-in for the induction surface of ``pyvl.cvl.Mesh`` backed by libpvl_hip.so.

Method names, positional signatures, validation order and exception types follow the
reference CPython type (src/meshobject.c:384-556, :956-1128, :1232-1286; stubs
python/pyvl/cvl.pyi:88-269; array checks src/common.h:8-66).  The numeric work runs on
the GPU:

- numpy arguments are staged through device memory and the result is written into a
  numpy array, exactly as the reference returns them;
- :class:`~pyvl_amd.device.DeviceArray` arguments stay resident and the result is a
  DeviceArray (no host round trip).

The optional ``line_buffer`` argument is validated as in the reference but is never
needed: the assembly kernels reduce lines into surfaces in registers.
"""

from __future__ import annotations

import ctypes as C
import operator
from typing import Sequence

import numpy as np

from . import _lib
from .device import DeviceArray

INVALID_ID = 0x7FFFFFFF  # src/core/common.h:54
_REVERSED = 0x80000000


# ---------------------------------------------------------------------------------------
# small value types used by Mesh.get_line / Mesh.get_surface (src/lineobject.c,
# src/surfaceobject.c:215-236)
class Line:
    """A line between two mesh points."""

    __slots__ = ("begin", "end")

    def __init__(self, begin: int, end: int) -> None:
        self.begin = int(begin)
        self.end = int(end)

    def __eq__(self, other: object) -> bool:
        return isinstance(other, Line) and self.begin == other.begin and self.end == other.end

    def __repr__(self) -> str:
        return f"Line({self.begin}, {self.end})"


class Surface:
    """A closed chain of lines."""

    __slots__ = ("_lines",)

    def __init__(self, lines: Sequence[Line]) -> None:
        self._lines = tuple(lines)

    @property
    def lines(self) -> tuple[Line, ...]:
        return self._lines

    @property
    def n_lines(self) -> int:
        return len(self._lines)


# ---------------------------------------------------------------------------------------
# argument checks (src/common.h:8-66)
def _ensure_array(arr, ndim: int, dims, name: str, *, writeable: bool = False) -> np.ndarray:
    if not isinstance(arr, np.ndarray):
        raise TypeError(f"{name} was not an array, but was instead {type(arr)!r}")
    if ndim and arr.ndim != ndim:
        raise ValueError(
            f"{name} did not have the expected number of axis ({ndim} required, {arr.ndim} found)."
        )
    for i, d in enumerate(dims):
        if d != 0 and arr.shape[i] != d:
            raise ValueError(
                f"{name} did not have the expected shape (axis {i} should have the size of {d}, "
                f"but was instead {arr.shape[i]})."
            )
    if not arr.flags.c_contiguous:
        raise ValueError(f"{name} was not C-contiguous.")
    if writeable and not arr.flags.writeable:
        raise ValueError(f"{name} was not writable.")
    if not arr.flags.aligned:
        raise ValueError(f"{name} was not aligned.")
    if arr.dtype != np.float64:
        raise ValueError(f"{name} did not have the correct data type.")
    return arr


def _ensure_device(arr, ndim: int, dims, name: str) -> DeviceArray:
    if not isinstance(arr, DeviceArray):
        raise TypeError(f"{name} must be a DeviceArray when the other inputs are device-resident")
    if arr.ndim != ndim:
        raise ValueError(
            f"{name} did not have the expected number of axis ({ndim} required, {arr.ndim} found)."
        )
    for i, d in enumerate(dims):
        if d != 0 and arr.shape[i] != d:
            raise ValueError(
                f"{name} did not have the expected shape (axis {i} should have the size of {d}, "
                f"but was instead {arr.shape[i]})."
            )
    if arr.dtype != np.float64:
        raise ValueError(f"{name} did not have the correct data type.")
    return arr


def _as_tol(x) -> float:
    """PyFloat_AsDouble semantics: numbers with __float__/__index__, not strings."""
    if isinstance(x, (str, bytes, bytearray)):
        raise TypeError(f"must be real number, not {type(x).__name__}")
    try:
        return float(x)
    except TypeError:
        return float(operator.index(x))


def _line_buffer_check(buf, n_lines: int, n_cpts: int) -> None:
    # src/meshobject.c:353-382
    if not isinstance(buf, np.ndarray):
        raise TypeError("Line computation buffer is not an array.")
    if buf.dtype != np.float64:
        raise ValueError("Line computation buffer was not an array of numpy.float64.")
    if not (buf.flags.c_contiguous and buf.flags.aligned and buf.flags.writeable):
        raise ValueError("Line computation buffer was not writable, C-contiguous, and aligned.")
    if buf.size < n_lines * n_cpts * 3:
        raise ValueError(
            "Line computation buffer did not have space for enough elements "
            f"(required {n_lines * n_cpts * 3}, but got {buf.size})."
        )


def _hp(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


class Mesh:
    """Mesh topology (points, lines, surfaces) with GPU induction methods.

    Mirrors ``pyvl.cvl.Mesh``.  Construction follows src/meshobject.c:25-124: the
    connectivity is a sequence of element index sequences of length >= 3, and an index
    equal to ``n_points`` is accepted as in the reference (its check is ``v > n_points``);
    such a mesh can still be inspected, but induction calls reject it because the point
    would be read out of bounds.
    """

    __slots__ = ("_n_points", "_lines", "_offsets", "_entries", "_dev", "_dev_lines", "__weakref__")

    def __new__(cls, n_points: int, connectivity):
        self = object.__new__(cls)
        self._dev = None
        self._dev_lines = None
        n_points = int(n_points)
        if n_points < 0 or n_points > 0xFFFFFFFF:
            raise OverflowError("n_points out of range for an unsigned int")
        try:
            seq = list(connectivity)
        except TypeError:
            raise TypeError("Second parameter must be a sequence of sequences") from None
        counts = np.empty(len(seq), np.uint32)
        flats = []
        for i, e in enumerate(seq):
            try:
                ln = len(e)
            except TypeError:
                raise TypeError(f"Element indices for element {i} were not a sequence.") from None
            if ln < 3:
                raise ValueError(f"Element {i} had only {ln} indices given (at least 3 are needed).")
            idx = np.asarray(e).astype(np.uint32, casting="unsafe").reshape(-1)
            bad = np.nonzero(idx > n_points)[0]
            if bad.size:
                k = int(bad[0])
                raise ValueError(
                    f"Element {i} had specified a point with index {int(idx[k])} as its {k} point, "
                    f"while only {n_points} points were given."
                )
            counts[i] = ln
            flats.append(idx)
        flat = np.concatenate(flats) if flats else np.empty(0, np.uint32)
        self._n_points = n_points
        self._build_from_elements(counts, flat)
        return self

    # -- construction helpers -----------------------------------------------------------
    def _build_from_elements(self, counts: np.ndarray, flat: np.ndarray) -> None:
        counts = np.ascontiguousarray(counts, np.uint32)
        flat = np.ascontiguousarray(flat, np.uint32)
        total = int(flat.size)
        lines = np.empty((max(total, 1), 2), np.uint32)
        offsets = np.empty(counts.size + 1, np.uint32)
        entries = np.empty(max(total, 1), np.uint32)
        n_lines = C.c_uint()
        _lib.call(
            "pvl_mesh_from_elements",
            C.c_uint(counts.size),
            counts.ctypes.data_as(C.POINTER(C.c_uint)),
            flat.ctypes.data_as(C.POINTER(C.c_uint)),
            C.byref(n_lines),
            lines.ctypes.data_as(C.POINTER(C.c_uint32)),
            offsets.ctypes.data_as(C.POINTER(C.c_uint32)),
            entries.ctypes.data_as(C.POINTER(C.c_uint32)),
        )
        self._lines = np.ascontiguousarray(lines[: n_lines.value])
        self._offsets = offsets
        self._entries = np.ascontiguousarray(entries[:total])

    @classmethod
    def _from_arrays(cls, n_points: int, lines, offsets, entries) -> "Mesh":
        self = object.__new__(cls)
        self._dev = None
        self._dev_lines = None
        self._n_points = int(n_points)
        self._lines = np.ascontiguousarray(lines, np.uint32).reshape(-1, 2)
        self._offsets = np.ascontiguousarray(offsets, np.uint32)
        self._entries = np.ascontiguousarray(entries, np.uint32)
        return self

    @classmethod
    def wake_rings(cls, n_strips: int, line_rows: int) -> "Mesh":
        """Ring mesh of ``WakeModelLineExplicitUnsteady._create_wake_mesh`` in closed form.

        Identical to ``Mesh(2*(R+1)*n, rings)`` with the ring list of
        python/pyvl/wake_models.py:35-52.
        """
        n_lines_total = n_strips * (3 * line_rows + 1)
        lines = np.empty((max(n_lines_total, 1), 2), np.uint32)
        offsets = np.empty(n_strips * line_rows + 1, np.uint32)
        entries = np.empty(max(4 * n_strips * line_rows, 1), np.uint32)
        n_lines = C.c_uint()
        _lib.call(
            "pvl_wake_topology",
            C.c_uint(n_strips),
            C.c_uint(line_rows),
            C.byref(n_lines),
            lines.ctypes.data_as(C.POINTER(C.c_uint32)),
            offsets.ctypes.data_as(C.POINTER(C.c_uint32)),
            entries.ctypes.data_as(C.POINTER(C.c_uint32)),
        )
        return cls._from_arrays(
            2 * (line_rows + 1) * n_strips,
            lines[: n_lines.value],
            offsets,
            entries[: 4 * n_strips * line_rows],
        )

    # -- properties -----------------------------------------------------------------------
    @property
    def n_points(self) -> int:
        return self._n_points

    @property
    def n_lines(self) -> int:
        return int(self._lines.shape[0])

    @property
    def n_surfaces(self) -> int:
        return int(self._offsets.size - 1)

    @property
    def line_data(self) -> np.ndarray:
        """Line connectivity (n_lines, 2) uint32 (a read-only view)."""
        v = self._lines.view()
        v.flags.writeable = False
        return v

    @property
    def surface_offsets(self) -> np.ndarray:
        return self._offsets

    @property
    def surface_lines(self) -> np.ndarray:
        return self._entries

    def __str__(self) -> str:
        return f"Mesh({self.n_points} points, {self.n_lines} lines, {self.n_surfaces} surfaces)"

    __repr__ = __str__

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, Mesh):
            return NotImplemented
        return (
            self.n_points == other.n_points
            and np.array_equal(self._lines, other._lines)
            and np.array_equal(self._offsets, other._offsets)
            and np.array_equal(self._entries, other._entries)
        )

    # -- element access -------------------------------------------------------------------
    def get_line(self, i) -> Line:
        idx = operator.index(i)
        n = self.n_lines
        if idx >= n or idx < -n:
            raise IndexError(f"Index {idx} is our of bounds for a mesh with {n} lines.")
        if idx < 0:
            idx += n
        a, b = self._lines[idx]
        return Line(a & INVALID_ID, b & INVALID_ID)

    def get_surface(self, i) -> Surface:
        idx = operator.index(i)
        n = self.n_surfaces
        if idx >= n or idx < -n:
            raise IndexError(f"Index {idx} is our of bounds for a mesh with {n} surfaces.")
        if idx < 0:
            idx += n
        out = []
        for e in self._entries[self._offsets[idx] : self._offsets[idx + 1]]:
            a, b = self._lines[e & INVALID_ID]
            # src/surfaceobject.c:223-233
            if e & _REVERSED:
                out.append(Line(a & INVALID_ID, b & INVALID_ID))
            else:
                out.append(Line(b & INVALID_ID, a & INVALID_ID))
        return Surface(out)

    def to_element_connectivity(self) -> tuple[np.ndarray, np.ndarray]:
        """Point counts and flat point list (src/core/mesh.c:249-301)."""
        counts = np.diff(self._offsets).astype(np.uintc)
        ent = self._entries
        rev = (ent & _REVERSED) != 0
        ln = self._lines[ent & INVALID_ID]
        flat = np.where(rev, ln[:, 0], ln[:, 1]) & INVALID_ID
        return counts, flat.astype(np.uintc)

    def compute_dual(self) -> "Mesh":
        """Dual mesh (src/core/mesh.c:70-173)."""
        nl = self.n_lines
        dual_lines = np.empty((max(nl, 1), 2), np.uint32)
        dual_off = np.empty(self._n_points + 1, np.uint32)
        dual_ent = np.empty(max(2 * nl, 1), np.uint32)
        _lib.call(
            "pvl_mesh_dual",
            C.c_uint(self._n_points),
            C.c_uint(nl),
            self._lines.ctypes.data_as(C.POINTER(C.c_uint32)),
            C.c_uint(self.n_surfaces),
            self._offsets.ctypes.data_as(C.POINTER(C.c_uint32)),
            self._entries.ctypes.data_as(C.POINTER(C.c_uint32)),
            dual_lines.ctypes.data_as(C.POINTER(C.c_uint32)),
            dual_off.ctypes.data_as(C.POINTER(C.c_uint32)),
            dual_ent.ctypes.data_as(C.POINTER(C.c_uint32)),
        )
        return Mesh._from_arrays(self.n_surfaces, dual_lines[:nl], dual_off, dual_ent[: int(dual_off[-1])])

    @classmethod
    def merge_meshes(cls, *meshes: "Mesh") -> "Mesh":
        """Concatenate meshes with index offsets (src/meshobject.c:780-860)."""
        for i, m in enumerate(meshes):
            if not isinstance(m, Mesh):
                raise TypeError(f"Element {i} in the input sequence was not a Mesh, but was instead {type(m)!r}")
        lines, offs, ents = [], [np.zeros(1, np.uint32)], []
        pts = lns = entr = 0
        for m in meshes:
            lv = m._lines.astype(np.uint64)
            val = (lv & INVALID_ID) + pts
            lines.append(((lv & _REVERSED) | val).astype(np.uint32))
            ev = m._entries.astype(np.uint64)
            ents.append(((ev & _REVERSED) | ((ev & INVALID_ID) + lns)).astype(np.uint32))
            offs.append((m._offsets[1:].astype(np.uint64) + entr).astype(np.uint32))
            pts += m.n_points
            lns += m.n_lines
            entr += int(m._offsets[-1])
        return cls._from_arrays(
            pts,
            np.concatenate(lines) if lines else np.empty((0, 2), np.uint32),
            np.concatenate(offs),
            np.concatenate(ents) if ents else np.empty(0, np.uint32),
        )

    def dual_normal_criterion(self, crit, normals) -> np.ndarray:
        """Dual lines whose end-point normals have dot product < crit (src/meshobject.c:1064-1128)."""
        c = _as_tol(crit)
        if c > 1.0 or c < -1.0:
            raise ValueError(
                f"Dot product criterion was {c:g}, which is not inside the allowed range of -1.0 to +1.0."
            )
        nrm = _ensure_array(normals, 2, (self._n_points, 3), "Normals array")
        ln = self._lines
        a, b = ln[:, 0] & INVALID_ID, ln[:, 1] & INVALID_ID
        ok = (a != INVALID_ID) & (b != INVALID_ID)
        idx = np.nonzero(ok)[0]
        na, nb = nrm[a[idx]], nrm[b[idx]]
        dp = na[:, 0] * nb[:, 0] + na[:, 1] * nb[:, 1] + na[:, 2] * nb[:, 2]
        return idx[dp < c].astype(np.uintc)

    def dual_free_edges(self) -> np.ndarray:
        ln = self._lines
        bad = ((ln[:, 0] & INVALID_ID) == INVALID_ID) | ((ln[:, 1] & INVALID_ID) == INVALID_ID)
        return np.nonzero(bad)[0].astype(np.uintc)

    def line_velocities_from_point_velocities(self, point_velocities, out) -> None:
        """Line velocity = mean of end-point velocities (src/meshobject.c:739-778)."""
        v = _ensure_array(point_velocities, 2, (self._n_points, 3), "Point velocities")
        o = _ensure_array(out, 2, (self.n_lines, 3), "Output array", writeable=True)
        a, b = self._lines[:, 0] & INVALID_ID, self._lines[:, 1] & INVALID_ID
        o[...] = (v[a] + v[b]) * 0.5

    # -- device mesh ----------------------------------------------------------------------
    def _device_lines(self) -> DeviceArray:
        """The line connectivity alone in device memory (uint32 pairs, INVALID_ID allowed),
        uploaded once: what Mesh.line_forces reads of a dual mesh, which pvl_mesh_create
        (points must exist) would not accept."""
        d = self._dev_lines
        if d is None:
            ld = np.ascontiguousarray(self.line_data.reshape(-1), np.uint32)
            d = DeviceArray((max(ld.size, 1),), np.uint32)
            if ld.size:
                d.copy_from(ld)
            self._dev_lines = d
        return d

    def _device(self):
        if self._dev is None:
            h = C.c_void_p()
            lines = np.ascontiguousarray(self._lines, np.uint32)
            _lib.call(
                "pvl_mesh_create",
                C.byref(h),
                C.c_uint(self._n_points),
                C.c_uint(self.n_lines),
                lines.ctypes.data_as(C.POINTER(C.c_uint32)),
                C.c_uint(self.n_surfaces),
                self._offsets.ctypes.data_as(C.POINTER(C.c_uint32)),
                self._entries.ctypes.data_as(C.POINTER(C.c_uint32)),
            )
            self._dev = h
        return self._dev

    def __del__(self):
        try:
            if self._dev is not None and _lib.alive():
                _lib._LIB.pvl_mesh_destroy(self._dev)
                self._dev = None
        except Exception:
            pass

    # -- geometry -------------------------------------------------------------------------
    def _surface_vec3(self, fn: str, args) -> np.ndarray | DeviceArray:
        if len(args) not in (1, 2):
            raise TypeError(
                f"Function must be called with either 1 or 2 arguments, but {len(args)} were given."
            )
        pos = args[0]
        out = args[1] if len(args) == 2 else None
        if isinstance(pos, DeviceArray):
            _ensure_device(pos, 2, (self._n_points, 3), "Input array")
            if out is None:
                out = DeviceArray((self.n_surfaces, 3))
            _ensure_device(out, 2, (self.n_surfaces, 3), "Output array")
            _lib.call(fn, self._device(), pos.cptr, out.cptr)
            return out
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64))
        if p.ndim != 2:
            raise ValueError(f"object of too small depth for desired array")
        if p.shape[1] != 3 or p.shape[0] != self._n_points:
            raise ValueError(
                "Input array did not have the shape expected from the number of points in the mesh "
                f"(expected a ({self._n_points}, 3) array, but got ({p.shape[0]}, {p.shape[1]}))."
            )
        if out is None:
            out = np.empty((self.n_surfaces, 3), np.float64)
        else:
            _ensure_array(out, 2, (self.n_surfaces, 3), "Output array", writeable=True)
        dp = DeviceArray.from_numpy(p)
        do = DeviceArray((self.n_surfaces, 3))
        _lib.call(fn, self._device(), dp.cptr, do.cptr)
        do.numpy(out)
        return out

    def surface_normal(self, *args):
        """Unit surface normals (src/core/mesh.c:48-68)."""
        return self._surface_vec3("pvl_surface_normal", args)

    def surface_average_vec3(self, *args):
        """Mean of surface start points (src/core/mesh.c:24-46)."""
        return self._surface_vec3("pvl_surface_center", args)

    # -- induction ------------------------------------------------------------------------
    def induction_matrix3(self, *args, precision: int = _lib.PVL_FP64):
        """Normal induction (system) matrix ``(n_cpts, n_surfaces)``.

        ``induction_matrix3(tol, positions, control_points, normals, out=None, line_buffer=None)``
        (src/meshobject.c:384-473).
        """
        nargs = len(args)
        if nargs not in (4, 5, 6):
            raise TypeError(f"Method requires 4, 5, or 6 arguments, but was called with {nargs}.")
        tol = _as_tol(args[0])
        pos, cpts, nrm = args[1], args[2], args[3]
        out = args[4] if nargs > 4 else None
        lbuf = args[5] if nargs == 6 else None
        ns = self.n_surfaces
        if isinstance(pos, DeviceArray):
            _ensure_device(pos, 2, (self._n_points, 3), "Position array")
            _ensure_device(cpts, 2, (0, 3), "Control point array")
            m = cpts.shape[0]
            _ensure_device(nrm, 2, cpts.shape, "Normal array")
            if out is None:
                out = DeviceArray((m, ns))
            _ensure_device(out, 2, (m, ns), "Output array")
            _lib.call(
                "pvl_induction_matrix3", self._device(), C.c_double(tol), pos.cptr, C.c_uint(m), cpts.cptr,
                nrm.cptr, out.cptr, C.c_int(precision),
            )
            return out
        _ensure_array(pos, 2, (self._n_points, 3), "Position array")
        _ensure_array(cpts, 2, (0, 3), "Control point array")
        m = cpts.shape[0]
        _ensure_array(nrm, cpts.ndim, cpts.shape, "Normal array")
        if out is not None:
            _ensure_array(out, 2, (m, ns), "Output array", writeable=True)
        else:
            out = np.empty((m, ns), np.float64)
        if lbuf is not None:
            _line_buffer_check(lbuf, self.n_lines, m)
        _lib.call(
            "pvl_induction_matrix3_host", self._device(), C.c_double(tol), _hp(pos), C.c_uint(m), _hp(cpts),
            _hp(nrm), _hp(out), C.c_int(precision),
        )
        return out

    def induction_matrix(self, *args, precision: int = _lib.PVL_FP64):
        """Surface induction tensor ``(n_cpts, n_surfaces, 3)``.

        ``induction_matrix(tol, positions, control_points, out=None, line_buffer=None)``
        (src/meshobject.c:475-556).
        """
        nargs = len(args)
        if nargs not in (3, 4, 5):
            raise TypeError(f"Method requires 3, 4, or 5 arguments, but was called with {nargs}.")
        tol = _as_tol(args[0])
        pos, cpts = args[1], args[2]
        out = args[3] if nargs > 3 else None
        lbuf = args[4] if nargs == 5 else None
        ns = self.n_surfaces
        if isinstance(pos, DeviceArray):
            _ensure_device(pos, 2, (self._n_points, 3), "Position")
            _ensure_device(cpts, 2, (0, 3), "Control point array")
            m = cpts.shape[0]
            if out is None:
                out = DeviceArray((m, ns, 3))
            _ensure_device(out, 3, (m, ns, 3), "Output tensor")
            _lib.call(
                "pvl_induction_matrix", self._device(), C.c_double(tol), pos.cptr, C.c_uint(m), cpts.cptr,
                out.cptr, C.c_int(precision),
            )
            return out
        _ensure_array(pos, 2, (self._n_points, 3), "Position")
        _ensure_array(cpts, 2, (0, 3), "Control point array")
        m = cpts.shape[0]
        if out is not None:
            _ensure_array(out, 3, (m, ns, 3), "Output tensor", writeable=True)
        else:
            out = np.empty((m, ns, 3), np.float64)
        if lbuf is not None:
            _line_buffer_check(lbuf, self.n_lines, m)
        _lib.call(
            "pvl_induction_matrix_host", self._device(), C.c_double(tol), _hp(pos), C.c_uint(m), _hp(cpts),
            _hp(out), C.c_int(precision),
        )
        return out

    def line_induction_matrix(self, *args, precision: int = _lib.PVL_FP64):
        """Per-line induction tensor ``(n_cpts, n_lines, 3)``.

        ``line_induction_matrix(tol, positions, control_points, out=None)``
        (src/meshobject.c:1232-1286).
        """
        nargs = len(args)
        if nargs not in (3, 4):
            raise TypeError(f"Method requires 3, or 4 arguments, but was called with {nargs}.")
        tol = _as_tol(args[0])
        pos, cpts = args[1], args[2]
        out = args[3] if nargs > 3 else None
        nl = self.n_lines
        if isinstance(pos, DeviceArray):
            _ensure_device(pos, 2, (self._n_points, 3), "Position array")
            _ensure_device(cpts, 2, (0, 3), "Control point array")
            m = cpts.shape[0]
            if out is None:
                out = DeviceArray((m, nl, 3))
            _ensure_device(out, 3, (m, nl, 3), "Output tensor")
            _lib.call(
                "pvl_line_induction_matrix", self._device(), C.c_double(tol), pos.cptr, C.c_uint(m), cpts.cptr,
                out.cptr, C.c_int(precision),
            )
            return out
        _ensure_array(pos, 2, (self._n_points, 3), "Position array")
        _ensure_array(cpts, 2, (0, 3), "Control point array")
        m = cpts.shape[0]
        if out is not None:
            _ensure_array(out, 3, (m, nl, 3), "Output tensor", writeable=True)
        else:
            out = np.empty((m, nl, 3), np.float64)
        _lib.call(
            "pvl_line_induction_matrix_host", self._device(), C.c_double(tol), _hp(pos), C.c_uint(m), _hp(cpts),
            _hp(out), C.c_int(precision),
        )
        return out

    def point_velocity(self, tol, positions, circulation, points, out=None, *, accumulate: bool = False,
                       precision: int = _lib.PVL_FP64):
        """Velocity induced at ``points`` by the mesh's surface circulations.

        Equals ``np.vecdot(self.induction_matrix(tol, positions, points), circulation[None, :, None], axis=1)``
        (python/pyvl/wake_models.py:125-126, :170-175; python/pyvl/solver.py:274-280)
        without materialising the tensor.
        """
        tol = _as_tol(tol)
        dev = isinstance(positions, DeviceArray)
        if dev:
            pos = _ensure_device(positions, 2, (self._n_points, 3), "Position array")
            circ = _ensure_device(circulation, 1, (self.n_surfaces,), "Circulation array")
            pts = _ensure_device(points, 2, (0, 3), "Point array")
            if out is None:
                out = DeviceArray.zeros(pts.shape) if accumulate else DeviceArray(pts.shape)
            _ensure_device(out, 2, pts.shape, "Output array")
            _lib.call(
                "pvl_point_velocity", self._device(), C.c_double(tol), pos.cptr, circ.cptr, C.c_uint(pts.shape[0]),
                pts.cptr, out.cptr, C.c_int(int(accumulate)), C.c_int(precision),
            )
            return out
        p = DeviceArray.from_numpy(np.asarray(positions, np.float64))
        c = DeviceArray.from_numpy(np.asarray(circulation, np.float64))
        q = DeviceArray.from_numpy(np.asarray(points, np.float64))
        o = DeviceArray.from_numpy(out) if (out is not None and accumulate) else None
        r = self.point_velocity(tol, p, c, q, o, accumulate=accumulate, precision=precision)
        if out is None:
            return r.numpy()
        r.numpy(out)
        return out

    @staticmethod
    def line_forces(primal: "Mesh", dual: "Mesh", circulation, positions, freestream, out=None):
        """Forces on the primal lines from their (reduced) line circulation in the mean velocity
        of their end points (src/meshobject.c:1288-1396): for line (p1, p2) whose dual line is
        (s+, s-), c = circulation[s+] - circulation[s-] (INVALID_ID ends skipped) and
        f = (r_p2 - r_p1) x ((v_p1 + v_p2) * (0.5 c)).  Static, as in the reference.

        numpy arguments: returns (n_lines, 3) numpy (or fills ``out``).  DeviceArray
        arguments: stays in HBM and returns a DeviceArray.
        """
        if not isinstance(primal, Mesh) or not isinstance(dual, Mesh):
            raise TypeError("primal and dual must be Mesh objects")
        n_lines = primal.n_lines
        if primal.n_points != dual.n_surfaces or n_lines != dual.n_lines or primal.n_surfaces != dual.n_points:
            raise ValueError(
                "Given meshes can not be dual to each other, since the number of points,lines, and surfaces don't "
                f"match as primal ({primal.n_points}, {n_lines}, {primal.n_surfaces}) and dual "
                f"({dual.n_points}, {dual.n_lines}, {dual.n_surfaces}).")
        if isinstance(positions, DeviceArray):
            c = _ensure_device(circulation, 1, (primal.n_surfaces,), "Circulation array")
            p = _ensure_device(positions, 2, (primal.n_points, 3), "Positions array")
            v = _ensure_device(freestream, 2, (primal.n_points, 3), "Velocity array")
            if out is None:
                out = DeviceArray((n_lines, 3))
            _ensure_device(out, 2, (n_lines, 3), "Output array")
            _lib.call("pvl_line_forces", primal._device(), dual._device_lines().cptr, C.c_uint(dual.n_points),
                      C.c_uint(dual.n_lines), C.c_uint(dual.n_surfaces), c.cptr, p.cptr, v.cptr, out.cptr)
            return out
        for a in (circulation, positions, freestream):
            if not isinstance(a, np.ndarray):
                raise TypeError(f"argument must be numpy.ndarray, not {type(a).__name__}")
        if out is not None and not isinstance(out, np.ndarray):
            raise TypeError(f"argument 'out' must be numpy.ndarray, not {type(out).__name__}")
        c = _ensure_array(circulation, 1, (primal.n_surfaces,), "Circulation array")
        p = _ensure_array(positions, 2, (primal.n_points, 3), "Positions array")
        v = _ensure_array(freestream, 2, (primal.n_points, 3), "Velocity array")
        if out is not None:
            _ensure_array(out, 2, (n_lines, 3), "Output array", writeable=True)
        r = Mesh.line_forces(primal, dual, DeviceArray.from_numpy(c), DeviceArray.from_numpy(p),
                             DeviceArray.from_numpy(v))
        if out is None:
            return r.numpy()
        r.numpy(out)
        return out
