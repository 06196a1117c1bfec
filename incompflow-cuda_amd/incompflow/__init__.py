"""MI355X steady SIMPLE solver: Python binding of the C-ABI in include/incompflow.h."""
from . import _abi  # noqa: F401
from ._abi import (Config, IterRecord, Report, default_config, channel_config, cavity_config,  # noqa: F401
                   cavity3d_config, adjoint_channel_config, read_state, write_state, StateHeader)
from ._lib import Solver, Batch, IflError, lib, LIB_PATH, EXPORTS, solve_noalloc  # noqa: F401
