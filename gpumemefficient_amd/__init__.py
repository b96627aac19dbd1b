"""gpumemefficient_amd -- MI355X (gfx950) wavefront DP engine for LCS, edit
distance and Smith-Waterman, plus the in-place SAT and 2D-SOR table sweeps
that share its wavefront schedule.

The compute path is libwfdp.so (hand-written HIP, C ABI in include/wfdp.h);
this package is the Python host side over it.
"""
from .api import ED, LCS, SOR, SW, Context, Job, Plan, Result, Tile, WfError, default_context, tile_for, variants

__all__ = ["Context", "Plan", "Job", "Tile", "Result", "WfError", "LCS", "SW", "ED", "SOR", "default_context",
           "variants", "tile_for"]
