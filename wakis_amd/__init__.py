"""wakis_amd: MI355X-native FIT time-stepping for Wakis-style wakefield and
impedance simulations.

Public names follow the reference package (wakis/__init__.py:6-24):
``Field, GridFIT3D, SolverFIT3D, WakeSolver`` plus the ``sources`` and
``materials`` modules.  The time stepping, bunch injection, wake probe,
wake potential and impedance DFT run in hand-written HIP kernels
(``wakis_amd/csrc``) behind the C ABI of ``include/wakis_fit.h``.
"""

from . import field, field_monitors, materials, sources  # noqa: F401
from .field import Field  # noqa: F401
from .field_monitors import FieldMonitor  # noqa: F401
from .grid import GridFIT3D  # noqa: F401
from .solver import SolverFIT3D  # noqa: F401
from .sources import Beam, PlaneWave  # noqa: F401
from .wake import WakeSolver  # noqa: F401
from .batch import BatchStepper, wakesolve_batch  # noqa: F401

__version__ = "0.1.0"
