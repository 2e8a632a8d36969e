"""MI355X-native CAC force-and-neighbor path.

The compute path is libcac_hip.so (hand-written HIP for gfx950 behind the
C-ABI in include/cac_hip.h); this package builds it, binds it with ctypes and
mirrors the reference's pair-style / run interface on top of it.
"""
from .engine import Engine, PairEAM, PairLJCut, PairError, eam_tables  # noqa: F401
from .lattice import CACSystem, build_cac_system, corner_quadrature  # noqa: F401
from .potential import Funcfl, johnson_cu_funcfl, read_funcfl, write_funcfl, lj_coeffs  # noqa: F401
from .datafile import DataFile, read_data, write_data  # noqa: F401

__all__ = ["Engine", "PairEAM", "PairLJCut", "PairError", "eam_tables", "CACSystem", "build_cac_system",
           "corner_quadrature", "Funcfl", "johnson_cu_funcfl", "read_funcfl", "write_funcfl", "lj_coeffs",
           "DataFile", "read_data", "write_data"]
