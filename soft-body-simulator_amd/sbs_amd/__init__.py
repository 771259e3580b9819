"""Python plumbing over the MI355X XPBD library (include/sbs_gpu.h, libsbs_gpu.so)."""
from ._lib import (KERNEL_NAMES, LIB_PATH, WARM_START, Mlp, SbsError, Simulation, environment_domain, exclusive_scan, extract_boundary, grid_from_mesh,
                   grid_node_count, grid_node_positions, sort_pairs, bar5, boundary, box6, ellipsoid6, header_functions, lib,
                   lumped_mass)
from . import ply, scene_file, scenes

__all__ = ["KERNEL_NAMES", "LIB_PATH", "WARM_START", "Mlp", "SbsError", "Simulation", "environment_domain", "exclusive_scan", "extract_boundary",
           "grid_from_mesh", "grid_node_count", "grid_node_positions", "sort_pairs", "bar5", "boundary", "box6", "ellipsoid6", "header_functions",
           "lib", "lumped_mass", "ply", "scene_file", "scenes"]
