"""MI355X (gfx950) implementation of tess2's post-tessellation path.

Circumcenters, Voronoi cell volumes and the Voronoi density estimator on a
regular grid, behind the C ABI of include/tess2_amd.h (libtess2_amd.so).
"""
from .host import (DENSE_CIC, DENSE_TESS, EXACT_VORONOI, FP64, MIXED, REFERENCE_SCAN,
                    DenseResult, StagedRun, Tess2Error, block_grid_params, dense,
                    fill_circumcenters, grid_step_params, volume, volumes, write_grid)
from .ingest import HostBlock, make_block, periodic_block, periodic_ghosts, vert_to_tet

__all__ = ["DENSE_CIC", "DENSE_TESS", "EXACT_VORONOI", "FP64", "MIXED", "REFERENCE_SCAN",
           "DenseResult", "StagedRun", "Tess2Error", "block_grid_params", "dense",
           "fill_circumcenters", "grid_step_params", "volume", "volumes", "write_grid",
           "HostBlock", "make_block", "periodic_block", "periodic_ghosts", "vert_to_tet"]
