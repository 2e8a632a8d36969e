"""MI355X-native CAC force-and-neighbor path.

The compute path is libcac_hip.so (hand-written HIP for gfx950 behind the
C-ABI in include/cac_hip.h, with the native file readers and writers of
include/cac_io.h); this package builds it, binds it with ctypes and mirrors
the reference's pair-style / run interface on top of it.
"""
from .engine import Engine, PairEAM, PairEAMAlloy, PairLJCut, PairError, eam_tables, eam_tables_setfl  # noqa: F401
from .lattice import CACSystem, build_cac_system, corner_quadrature  # noqa: F401
from .potential import (Funcfl, Setfl, johnson_cu_funcfl, read_funcfl, write_funcfl, read_setfl,  # noqa: F401
                        write_setfl, setfl_from_funcfl, lj_coeffs)
from .datafile import DataFile, read_data, write_data  # noqa: F401
from .dump import write_dump_atom, write_dump_element, read_dump_atom, read_dump_element  # noqa: F401

__all__ = ["Engine", "PairEAM", "PairEAMAlloy", "PairLJCut", "PairError", "eam_tables", "eam_tables_setfl",
           "CACSystem", "build_cac_system", "corner_quadrature", "Funcfl", "Setfl", "johnson_cu_funcfl",
           "read_funcfl", "write_funcfl", "read_setfl", "write_setfl", "setfl_from_funcfl", "lj_coeffs",
           "DataFile", "read_data", "write_data", "write_dump_atom", "write_dump_element", "read_dump_atom",
           "read_dump_element"]
