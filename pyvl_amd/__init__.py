"""pyvl_amd: PyVL's line-vortex hot path on AMD Instinct MI355X (gfx950).

The numeric work lives in ``libpvl_hip.so`` (hand-written HIP kernels behind the C ABI in
``include/pvl.h``); this package is the host-side mirror of the reference's call surface:

- :class:`pyvl_amd.cvl.Mesh` -- ``pyvl.cvl.Mesh`` induction/geometry methods;
- :mod:`pyvl_amd.linalg` -- ``scipy.linalg.lu_factor``/``lu_solve`` as used by
  ``run_solver``, on the GPU;
- :mod:`pyvl_amd.wake_models` -- ``WakeModelLineExplicitUnsteady`` with its state resident
  in HBM;
- :mod:`pyvl_amd.solver` -- the ``run_solver`` step order.
"""

from .cvl import INVALID_ID, Line, Mesh, Surface
from .device import DeviceArray, synchronize

__all__ = ["INVALID_ID", "Line", "Mesh", "Surface", "DeviceArray", "synchronize"]
