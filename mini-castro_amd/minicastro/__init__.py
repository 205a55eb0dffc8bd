"""mini-castro-amd: mini-Castro's CTU hydro update on MI355X (gfx950).

The compute path is the HIP library ``lib/libmccastro.so`` built from
``csrc/``; this package binds its C ABI (``include/mc_castro.h``) with
ctypes and adds the problem configurations.  There is no CPU fallback: a
missing library raises.
"""
from ._lib import lib, lib_path, build, load_table, newton_failures, table_path  # noqa: F401
from .level import Level, MODE_PLAIN, MODE_FUSED, BC_PERIODIC, BC_OUTFLOW  # noqa: F401
from .ensemble import Ensemble  # noqa: F401
from .adjoint import Adjoint  # noqa: F401
from .batch import advance_batch, work_item  # noqa: F401
from . import amr, configs  # noqa: F401
