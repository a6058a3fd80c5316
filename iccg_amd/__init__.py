"""iccg_amd -- ICCG solve path for AMD MI355X (gfx950).

The compute lives in libiccg_hip.so (hand-written HIP kernels behind a C
ABI, include/iccg_hip.h); this package mirrors the stomo0511/ICCG interface
on top of it.
"""
from .api import (CRS, CGResult, Batch, ColorSchedule, DeviceCSR, IC0, Identity, Jacobi,
                  MultiSolver, Server, Solver, conjugate_gradient_multi,
                  BuildABMCSchedule, BuildColorPtrRows_FromPermutationAndColor,
                  BuildPermutationByBlockColor, BuildPermutationByColor,
                  Permute_PtAP_LowerCRS_to_LowerCRS, ReadBlockColorFile_1Based,
                  ReadBlockFile_1Based, ReadColorFile_1Based, abmc_blocking,
                  abmc_blocking_grid, bench_copy, check_hazards, check_interblock,
                  conjugate_gradient, coo_to_crs, device_count, device_trim, expand_lower_to_full,
                  gen_grid, gen_normal, greedy_color, read_matrix_market_crs,
                  write_block_color_file, write_block_file, write_color_file)
from ._capi import IccgError

__all__ = [n for n in dir() if not n.startswith("_")]
