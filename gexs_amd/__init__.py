"""gexs_amd -- MI355X (gfx950) line-by-line molecular cross-sections.

The compute path is libgexs_hip.so (hand-written HIP kernels behind the C ABI
of include/gexs_hip.h).  ``HipManager`` mirrors the reference's BaseManager
plugin API over that ABI; ``synth`` makes the seeded benchmark line lists.
Importing the package does not load the library; the first manager does, and
raises if the library is missing (there is no CPU fallback).
"""
from ._lib import GexsError, load  # noqa: F401
from .manager import HipManager  # noqa: F401

__all__ = ["HipManager", "GexsError", "load"]
