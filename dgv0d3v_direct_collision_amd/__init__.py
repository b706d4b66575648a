"""MI355X-native direct DGV collision evaluator.

Drop-in for the collision_mod / spat_oper_mod / time_integr_mod path of
alexmalekseenko/DGV0D3V_direct_collision: hand-written gfx950 HIP kernels behind
the C ABI of ``include/dgv.h`` (libdgv_amd.so), with this package as the
host-side mirror of the reference's procedures (see ``collision.DGV``).
"""
from ._lib import (DGV_FP32_COMP, DGV_FP32_F64ACC, DGV_FP64, DGV_INT8_SPLIT, DGV_KERNEL_AUTO, DGV_KERNEL_PAIRBLOCK,  # noqa: F401
                   DGV_KERNEL_MFMA, DGV_KERNEL_RECORD, PRECISION_NAMES, DGVError, LIB_PATH, load)
from .collision import (DGV, DGVBatch, agen_pair, chunk_name, grid_desc, grid_geometry,  # noqa: F401
                        log_file_name, read_sol_file, sol_file_name, write_sol_file,
                        write_sol_init_file)
from .driver_io import L1Track, MacroLog, MomDerivLog  # noqa: F401
from .surrogate import Surrogate  # noqa: F401

__all__ = ["DGV", "DGVBatch", "Surrogate", "DGVError", "agen_pair", "grid_geometry", "chunk_name", "grid_desc", "load", "LIB_PATH",
           "DGV_FP64", "DGV_FP32_COMP", "DGV_FP32_F64ACC", "DGV_INT8_SPLIT", "PRECISION_NAMES", "DGV_KERNEL_AUTO", "DGV_KERNEL_PAIRBLOCK",
           "DGV_KERNEL_RECORD", "DGV_KERNEL_MFMA"]
