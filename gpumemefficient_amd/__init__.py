"""gpumemefficient_amd -- MI355X (gfx950) wavefront DP engine for LCS, edit
distance and Smith-Waterman.

The compute path is libwfdp.so (hand-written HIP, C ABI in include/wfdp.h);
this package is the Python host side over it.
"""
from .api import ED, LCS, SW, Context, Job, Plan, Result, Tile, WfError, default_context, tile_for, variants

__all__ = ["Context", "Plan", "Job", "Tile", "Result", "WfError", "LCS", "SW", "ED", "default_context", "variants",
           "tile_for"]
