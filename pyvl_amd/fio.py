"""Serialisation of solver state: ``HirearchicalMap`` and its JSON form
(python/pyvl/fio/io_common.py, io_json.py), plus an exact binary form.

The reference writes JSON or HDF5 (h5py).  h5py is not part of this stack, so the binary
form here is numpy's ``.npz`` (flattened keys ``a/b/c``, loaded with ``allow_pickle=False``);
like HDF5 it stores arrays bit for bit.  JSON also round-trips float64 exactly (shortest
repr), as the reference's encoder does.
"""

from __future__ import annotations

import json
from collections.abc import MutableMapping, Sequence
from pathlib import Path
from typing import Any

import numpy as np


class HirearchicalMap(MutableMapping):
    """A mapping of string keys to arrays, scalars, strings and nested maps
    (python/pyvl/fio/io_common.py:11-157): insert_* refuse to overwrite a key."""

    def __init__(self) -> None:
        self._map: dict[str, Any] = {}

    def __getitem__(self, key: str) -> Any:
        return self._map[key]

    def __setitem__(self, key: str, value: Any) -> None:
        self._map[key] = value

    def __delitem__(self, key: str) -> None:
        del self._map[key]

    def __iter__(self):
        return iter(self._map)

    def __len__(self) -> int:
        return len(self._map)

    def __contains__(self, key: object) -> bool:
        return key in self._map

    def _insert(self, key: str, value: Any) -> None:
        if not isinstance(key, str):
            raise TypeError(f"Key is not a string but a {type(key).__name__}.")
        if key in self._map:
            raise KeyError(f'Map already contains a key "{key}".')
        self._map[key] = value

    def insert_type(self, key: str, t: type) -> None:
        if not isinstance(t, type):
            raise TypeError(f"The value was not a type but {type(t).__name__}")
        self._insert(key, t.__module__ + "." + t.__name__)

    def insert_array(self, key: str, value) -> None:
        self._insert(key, np.array(value))

    def insert_string(self, key: str, value: str) -> None:
        if not isinstance(value, str):
            raise TypeError(f"The value was not a string but {type(value).__name__}")
        self._insert(key, value)

    def insert_scalar(self, key: str, value) -> None:
        if not isinstance(value, (int, float)):
            raise TypeError(f"The value was not an int or float but {type(value).__name__}")
        self._insert(key, value)

    def insert_int(self, key: str, value: int) -> None:
        if not isinstance(value, int):
            raise TypeError(f"The value was not an int but {type(value).__name__}")
        self._insert(key, value)

    def _recursion_check(self, value: "HirearchicalMap") -> bool:
        if value is self:
            return True
        return any(isinstance(v, HirearchicalMap) and v._recursion_check(value) for v in self._map.values())

    def insert_hirearchycal_map(self, key: str, value: "HirearchicalMap") -> None:
        if not isinstance(value, HirearchicalMap):
            raise TypeError(f"The value was not a HirearchicalMap but {type(value).__name__}")
        if self._recursion_check(value) or value._recursion_check(self):
            raise ValueError("Inserting the hierarchical map would cause cyclical hierarchy.")
        self._insert(key, value)

    def get_type(self, key: str) -> type:
        name = self[key]
        if not isinstance(name, str):
            raise TypeError(f"The value was not a type name but instead {type(name).__name__}")
        module_name, type_name = name.rsplit(".", 1)
        mod = __import__(module_name, fromlist=[type_name])
        cls = getattr(mod, type_name)
        if not isinstance(cls, type):
            raise TypeError(f"The value was not a type but {type(cls).__name__}")
        return cls

    def get_array(self, key: str) -> np.ndarray:
        return np.array(self[key])

    def get_string(self, key: str) -> str:
        v = self._map[key]
        if isinstance(v, np.ndarray) and v.ndim == 0:
            v = str(v)
        if not isinstance(v, str):
            raise TypeError(f"The value was not a string but {type(v).__name__}")
        return v

    def get_scalar(self, key: str):
        v = self._map[key]
        if isinstance(v, np.ndarray) and v.ndim == 0:
            v = v.item()
        if not isinstance(v, (int, float, np.integer, np.floating)):
            raise TypeError(f"The value was not an int or float but {type(v).__name__}")
        return v

    def get_int(self, key: str) -> int:
        v = self._map[key]
        if isinstance(v, np.ndarray) and v.ndim == 0:
            v = v.item()
        if not isinstance(v, (int, np.integer)):
            raise TypeError(f"The value was not an int but {type(v).__name__}")
        return int(v)

    def get_hirearchical_map(self, key: str) -> "HirearchicalMap":
        v = self._map[key]
        if not isinstance(v, HirearchicalMap):
            raise TypeError(f"The value was not a HirearchicalMap but {type(v).__name__}")
        return v

    @classmethod
    def from_dict(cls, d: dict) -> "HirearchicalMap":
        """Nested plain dicts (the save() maps of this package) as maps."""
        out = cls()
        for k, v in d.items():
            out[k] = cls.from_dict(v) if isinstance(v, dict) else v
        return out

    def to_dict(self) -> dict:
        return {k: (v.to_dict() if isinstance(v, HirearchicalMap) else v) for k, v in self._map.items()}


# -- JSON (python/pyvl/fio/io_json.py) -----------------------------------------------------
class _Encoder(json.JSONEncoder):
    def default(self, o: Any) -> Any:
        if isinstance(o, HirearchicalMap):
            return o._map
        if isinstance(o, np.ndarray):
            return o.tolist()
        if isinstance(o, np.integer):
            return int(o)
        if isinstance(o, np.floating):
            return float(o)
        return super().default(o)


def serialize_json(hmap: HirearchicalMap, path: Path | str) -> None:
    with open(path, "w") as f:
        f.write(_Encoder().encode(hmap))


def _hook(arg: dict) -> HirearchicalMap:
    hm = HirearchicalMap()
    for k, v in arg.items():
        hm[k] = np.array(v) if isinstance(v, Sequence) and not isinstance(v, str) else v
    return hm


def deserialize_json(path: Path | str) -> HirearchicalMap:
    with open(path) as f:
        return json.load(f, object_hook=_hook)


# -- exact binary form ---------------------------------------------------------------------
def _flatten(prefix: str, m: HirearchicalMap, out: dict) -> None:
    for k, v in m.items():
        if "/" in k:
            raise ValueError(f'key "{k}" contains "/"')
        name = f"{prefix}/{k}" if prefix else k
        if isinstance(v, HirearchicalMap):
            _flatten(name, v, out)
            if len(v) == 0:
                out[name + "/"] = np.zeros(0)
        else:
            out[name] = np.asarray(v)


def serialize_npz(hmap: HirearchicalMap, path: Path | str) -> None:
    flat: dict = {}
    _flatten("", hmap, flat)
    with open(path, "wb") as f:
        np.savez(f, **flat)


def deserialize_npz(path: Path | str) -> HirearchicalMap:
    root = HirearchicalMap()
    with np.load(path, allow_pickle=False) as z:
        for name in z.files:
            parts = name.split("/")
            node = root
            for p in parts[:-1]:
                if p not in node:
                    node[p] = HirearchicalMap()
                node = node[p]
            if parts[-1] == "":
                continue
            v = z[name]
            if v.ndim == 0:
                v = v.item() if v.dtype.kind in "iufb" else str(v)
            node[parts[-1]] = v
    return root
